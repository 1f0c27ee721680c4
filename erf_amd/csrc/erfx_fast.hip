// erfx_fast.hip -- acoustic substep on a flat grid: make_fast_coeffs and
// erf_fast_rhs_N as gfx950 kernels.
//
// Launch structure per substep:
//   K1  cells grown by 1: Delta(rho theta), extrapolated theta, lagged update
//   K2  x- and y-faces (one launch): horizontal momentum update, written
//       straight into the current state (each lane reads prev and writes cur
//       at its own index, so the S_prev == S_data aliasing is safe)
//   K3-5 columns (one lane per (i, j), k loops): horizontal divergence of the
//       increments, tridiagonal RHS, Thomas sweep, new rho*w, vertical
//       fluxes, time-averaged rho*w, rho and rho*theta update
// The arithmetic of every point keeps the reference's operation order.
#include "erfx_common.h"
#include "erfx_internal.h"
#include <mutex>
#include <string>
#include <type_traits>

namespace erfx {

// make_fast_coeffs, flat branch (Source/TimeIntegration/ERF_make_fast_coeffs.cpp:158-200),
// GPU column recurrence (:214-242) and inversion (:287-290), fused into one
// column kernel.  fc comps: 0 A, 1 B (1/B on interior faces), 2 C, 3 P, 4 Q.
void make_fast_coeffs(const Params& P, double* fc, const erfx_state& stage, const double* prim,
                      const double* pi_stage, const double* r0, const double* pi0, double dtau, double beta_s,
                      hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sz = L.sz;
    const double beta_2 = 0.5 * (1.0 + beta_s);
    const double c_v = P.cfg.c_p - R_d;
    const double dzi = P.dxInv[2];
    const double halfg = fabs(0.5 * (-P.gravity));
    const bool top_outflow = (P.cfg.bc[5] == ERFX_BC_OUTFLOW);
    const double* rt = stage.cons + ne;
    double* A = fc;
    double* B = fc + ne;
    double* Cc = fc + 2 * ne;
    double* Pc = fc + 3 * ne;
    double* Qc = fc + 4 * ne;
    const double D = dtau * dtau * beta_2 * beta_2 * dzi;
    const bool moist = P.moist;
    const double* pq1 = prim + PRIM_Q1 * ne;
    const double* pq2 = prim + PRIM_Q2 * ne;
    const double* J = P.terrain ? P.detJ : nullptr;
    launch_cols(0, nx - 1, 0, ny - 1, s, [=] __device__(int i, int j) {
        const long q0 = L.ix(i, j, 0);
        // bottom face, then the interior faces with the recurrence carried along
        A[q0] = 0.0; B[q0] = 1.0; Cc[q0] = 0.0;
        double bet = 1.0;
        double c_prev = 0.0;
        double r_lo = r0[q0], pib_lo = pi0[q0], pis_lo = pi_stage[q0], rt_lo = rt[q0];
        double pr_km2 = prim[q0 - sz], pr_km1 = prim[q0], pr_k = prim[q0 + sz];
        double q1_lo = moist ? pq1[q0] : 0.0, q2_lo = moist ? pq2[q0] : 0.0;
        struct Lv { double r, pib, pis, rt_, pkp1, q1, q2; };
        auto load = [&](long q, Lv& v) {
            v.r = r0[q]; v.pib = pi0[q]; v.pis = pi_stage[q]; v.rt_ = rt[q]; v.pkp1 = prim[q + sz];
            if (moist) { v.q1 = pq1[q]; v.q2 = pq2[q]; } else { v.q1 = 0.0; v.q2 = 0.0; }
        };
        constexpr int RD = 4;
        Lv ring[RD];
#pragma unroll
        for (int d = 0; d < RD; ++d)
            if (1 + d <= nz - 1) load(q0 + (1 + d) * sz, ring[d]);
        for (int kb = 1; kb <= nz - 1; kb += RD) {
#pragma unroll
            for (int d = 0; d < RD; ++d) {
                const int k = kb + d;
                if (k > nz - 1) break;
                const long q = q0 + k * sz;
                Lv v = ring[d];
                if (k + RD <= nz - 1) load(q + RD * sz, ring[d]);
                double rhobar_lo = r_lo, rhobar_hi = v.r;
                double pibar_lo = pib_lo, pibar_hi = v.pib;
                double pi_c = 0.5 * (pis_lo + v.pis);
                double coeff_P, coeff_Q, detJ_on_kface = 1.0;
                if (J) {   // terrain (ERF_make_fast_coeffs.cpp:113-137)
                    detJ_on_kface = 0.5 * (J[q] + J[q - sz]);
                    double inv_detJ_on_kface = 1. / detJ_on_kface;
                    coeff_P = -Gamma * R_d * dzi * inv_detJ_on_kface + halfg * R_d * rhobar_hi / (c_v * pibar_hi * v.rt_);
                    coeff_Q = Gamma * R_d * dzi * inv_detJ_on_kface + halfg * R_d * rhobar_lo / (c_v * pibar_lo * rt_lo);
                } else {
                    coeff_P = -Gamma * R_d * dzi + halfg * R_d * rhobar_hi / (c_v * pibar_hi * v.rt_);
                    coeff_Q = Gamma * R_d * dzi + halfg * R_d * rhobar_lo / (c_v * pibar_lo * rt_lo);
                }
                coeff_P *= pi_c;
                coeff_Q *= pi_c;
                Pc[q] = coeff_P;
                Qc[q] = coeff_Q;
                if (moist) {   // ERF_make_fast_coeffs.cpp:183-188
                    double qf = 0.5 * (v.q1 + q1_lo + v.q2 + q2_lo);
                    coeff_P /= (1.0 + qf);
                    coeff_Q /= (1.0 + qf);
                }
                double theta_t_lo = 0.5 * (pr_km2 + pr_km1);
                double theta_t_mid = 0.5 * (pr_km1 + pr_k);
                double theta_t_hi = 0.5 * (pr_k + v.pkp1);
                double a = D * (halfg - coeff_Q * theta_t_lo);
                double c = D * (-halfg + coeff_P * theta_t_hi);
                double b = (J ? detJ_on_kface : 1.0) + D * (coeff_Q - coeff_P) * theta_t_mid;
                A[q] = a;
                Cc[q] = c;
                double gam = c_prev / bet;
                bet = b - a * gam;
                c_prev = c;
                B[q] = 1.0 / bet;                        // interior faces store 1/B (:287-290)
                r_lo = v.r; pib_lo = v.pib; pis_lo = v.pis; rt_lo = v.rt_;
                pr_km2 = pr_km1; pr_km1 = pr_k; pr_k = v.pkp1;
                q1_lo = v.q1; q2_lo = v.q2;
            }
        }
        // top face keeps B itself
        const long qt = q0 + nz * sz;
        const double at = top_outflow ? -1.0 : 0.0;
        A[qt] = at; Cc[qt] = 0.0;
        double gam = c_prev / bet;
        B[qt] = 1.0 - at * gam;
    });
}

// no_substep_fun (Source/TimeIntegration/ERF_TI_no_substep_fun.H:88-126),
// fixed or no terrain: with substepping_type None each RK stage advances the
// fast variables with the slow right-hand side alone -- rho and rho*theta on
// the valid cells, each momentum on its nodal box -- in one launch.
void no_substep(const Params& P, const erfx_state& ssum, const erfx_state& sold, const erfx_state& fslow,
                double slow_dt, hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne;
    launch_box(Box(0, 0, 0, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (i <= nx - 1 && j <= ny - 1 && k <= nz - 1) {
            ssum.cons[q] = sold.cons[q] + slow_dt * (fslow.cons[q]);
            ssum.cons[q + ne] = sold.cons[q + ne] + slow_dt * (fslow.cons[q + ne]);
        }
        if (j <= ny - 1 && k <= nz - 1) ssum.xmom[q] = sold.xmom[q] + slow_dt * fslow.xmom[q];
        if (i <= nx - 1 && k <= nz - 1) ssum.ymom[q] = sold.ymom[q] + slow_dt * fslow.ymom[q];
        if (i <= nx - 1 && j <= ny - 1) ssum.zmom[q] = sold.zmom[q] + slow_dt * fslow.zmom[q];
    });
}

void fast_rhs_N(const Params& P, const FastWork& W, int step, int nrk, const erfx_state& slow_rhs,
                const erfx_state& prev, const erfx_state& stage, const double* prim, const double* pi_stage,
                const double* fc, const erfx_state& data, const erfx_state& scratch, double dtau,
                double beta_s, double facinv, hipStream_t s, const FastChain& ch) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const double beta_1 = 0.5 * (1.0 - beta_s);
    const double beta_2 = 0.5 * (1.0 + beta_s);
    const double beta_d = 0.1;
    const double dxi = P.dxInv[0], dyi = P.dxInv[1], dzi = P.dxInv[2];
    const bool implicit = P.cfg.implicit_substepping != 0;
    double* extrap = W.extrap;
    double* drw = W.drw;
    double* tr0 = W.trhs;
    double* tr1 = W.trhs + ne;
    double* soln = W.soln;
    const double* prt = prev.cons + ne;
    const double* srt = stage.cons + ne;
    double* lagged = scratch.cons + ne;

    // K1 (ERF_fast_rhs_N.cpp:115-152)
    launch_box(Box(-1, -1, -1, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        double d = prt[q] - srt[q];
        extrap[q] = (step == 0) ? d : d + beta_d * (d - lagged[q]);
        lagged[q] = d;
    });

    // K2 (:161-257).  With rt_prev the previous substep's apply_bcs is folded in:
    // its momentum -> velocity -> momentum round trip (ERF_MomentumToVelocity.cpp,
    // ERF_VelocityToMomentum.cpp) is applied to prev as it is read.
    const bool first = (nrk == 0 && step == 0);
    const bool rt = ch.rt_prev && step > 0;
    const double* prho = prev.cons;
    const double gr = -Gamma * R_d;
    const bool moist = P.moist;
    const double* pq1 = prim + PRIM_Q1 * ne;
    const double* pq2 = prim + PRIM_Q2 * ne;
    launch_box(Box(0, 0, 0, nx, ny, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (j <= ny - 1) {
            double pm = prev.xmom[q];
            if (rt) {
                double d = prho[q] + prho[q - 1];
                double u = pm * 2.0 / d;
                pm = u * 0.5 * d;
            }
            double new_drho_u;
            if (first) {
                new_drho_u = pm - stage.xmom[q] + dtau * slow_rhs.xmom[q];
            } else {
                double gpx = (extrap[q] - extrap[q - 1]) * dxi;
                if (moist) gpx /= (1.0 + face_q(pq1, pq2, q, q - 1));   // :215-219
                double pi_c = 0.5 * (pi_stage[q - 1] + pi_stage[q]);
                double fast_rhs_rho_u = gr * pi_c * gpx;
                new_drho_u = pm - stage.xmom[q] + dtau * fast_rhs_rho_u + dtau * slow_rhs.xmom[q];
            }
            scratch.xmom[q] += facinv * new_drho_u;
            data.xmom[q] = stage.xmom[q] + new_drho_u;
        }
        if (i <= nx - 1) {
            double pm = prev.ymom[q];
            if (rt) {
                double d = prho[q] + prho[q - sy];
                double v = pm * 2.0 / d;
                pm = v * 0.5 * d;
            }
            double new_drho_v;
            if (first) {
                new_drho_v = pm - stage.ymom[q] + dtau * slow_rhs.ymom[q];
            } else {
                double gpy = (extrap[q] - extrap[q - sy]) * dyi;
                if (moist) gpy /= (1.0 + face_q(pq1, pq2, q, q - sy));  // :238-242
                double pi_c = 0.5 * (pi_stage[q - sy] + pi_stage[q]);
                double fast_rhs_rho_v = gr * pi_c * gpy;
                new_drho_v = pm - stage.ymom[q] + dtau * fast_rhs_rho_v + dtau * slow_rhs.ymom[q];
            }
            scratch.ymom[q] += facinv * new_drho_v;
            data.ymom[q] = stage.ymom[q] + new_drho_v;
        }
    });

    // K3 + K4 + K5 fused per column (:328-357, :374-453, :518-536, :566-592).
    // One lane per (i, j) column.  Forward sweep: horizontal divergence of the
    // increments in cell k, tridiagonal RHS on face k, elimination.  Backward
    // sweep: substitution, new rho*w, vertical fluxes and the rho / rho*theta
    // update of cell k.  The loads of the next level are issued before the
    // arithmetic of the current one, so each lane keeps two levels of loads in
    // flight.  Every expression keeps the reference's operation order.
    const double halfg = fabs(0.5 * (-P.gravity));
    const double* A = fc;
    const double* iB = fc + ne;
    const double* Cc = fc + 2 * ne;
    const double* Pc = fc + 3 * ne;
    const double* Qc = fc + 4 * ne;
    const double* sr0 = slow_rhs.cons;
    const double* sr1 = slow_rhs.cons + ne;
    const double* pr0 = prev.cons;
    const double* st0 = stage.cons;
    double* cc0 = data.cons;
    double* cc1 = data.cons + ne;
    const bool gw = ch.ghost_writes;
    const double ch_w_lo = ch.w_lo, ch_w_hi = ch.w_hi;
    auto column = [=] __device__(int i, int j, auto depth) {
        const long q0 = L.ix(i, j, 0);
        // per-level loads of the forward sweep (face k and cell k)
        struct Fw {
            double P, Q, A, iB, ws, pkp1, pz, szz, pr, rkp1, st, prt_, srt_, s0, s1;
            double dxl, dxh, sxl, sxh, dyl, dyh, syl, syh, pc, pxm, pxp, pym, pyp;
            double q1, q2;   // qv, qc of cell k (moist only)
        };
        auto fload = [&](long q, Fw& v) {
            v.P = Pc[q]; v.Q = Qc[q]; v.A = A[q]; v.iB = iB[q]; v.ws = slow_rhs.zmom[q];
            v.pkp1 = prim[q + sz]; v.pz = prev.zmom[q + sz]; v.szz = stage.zmom[q + sz];
            v.pr = pr0[q]; v.rkp1 = rt ? pr0[q + sz] : 0.0;
            v.st = st0[q]; v.prt_ = prt[q]; v.srt_ = srt[q]; v.s0 = sr0[q]; v.s1 = sr1[q];
            v.dxl = data.xmom[q]; v.dxh = data.xmom[q + 1]; v.sxl = stage.xmom[q]; v.sxh = stage.xmom[q + 1];
            v.dyl = data.ymom[q]; v.dyh = data.ymom[q + sy]; v.syl = stage.ymom[q]; v.syh = stage.ymom[q + sy];
            v.pc = prim[q]; v.pxm = prim[q - 1]; v.pxp = prim[q + 1]; v.pym = prim[q - sy]; v.pyp = prim[q + sy];
            if (moist) { v.q1 = pq1[q]; v.q2 = pq2[q]; } else { v.q1 = 0.0; v.q2 = 0.0; }
        };
        // horizontal flux divergence of the increments (K3)
        auto hdiv = [&](const Fw& v, double& t0, double& t1) {
            double xflux_lo = v.dxl - v.sxl;
            double xflux_hi = v.dxh - v.sxh;
            double yflux_lo = v.dyl - v.syl;
            double yflux_hi = v.dyh - v.syh;
            double pc = v.pc;
            t0 = (xflux_hi - xflux_lo) * dxi + (yflux_hi - yflux_lo) * dyi;
            t1 = ((xflux_hi * (pc + v.pxp) - xflux_lo * (pc + v.pxm)) * dxi +
                  (yflux_hi * (pc + v.pyp) - yflux_lo * (pc + v.pym)) * dyi) * 0.5;
        };

        // prev rho*w at the face above a level's cell, after the round trip when
        // rt (interior: rho of the two cells; top: the Dirichlet w with the
        // foextrap ghost rho)
        auto w_above = [&](const Fw& v, int kf) {
            if (!rt) return v.pz;
            if (kf >= nz) return ch_w_hi * 0.5 * (v.pr + v.pr);
            double d = v.rkp1 + v.pr;
            double w = v.pz * 2.0 / d;
            return w * 0.5 * d;
        };
        Fw cur;
        fload(q0, cur);
        // face 0 (bottom row, :423-432) and cell 0
        double pz0 = rt ? ch_w_lo * 0.5 * (cur.pr + cur.pr) : prev.zmom[q0];
        double Om_km1 = pz0 - stage.zmom[q0];
        drw[q0] = Om_km1;
        double rhs0 = Om_km1 + dtau * cur.ws;
        double sol_prev = implicit ? rhs0 * cur.iB : rhs0;
        soln[q0] = sol_prev;
        double t0_km1, t1_km1;
        hdiv(cur, t0_km1, t1_km1);
        tr0[q0] = t0_km1;
        tr1[q0] = t1_km1;
        double drho_km1 = cur.pr - cur.st;
        double drt_km1 = cur.prt_ - cur.srt_;
        double s0_km1 = cur.s0, s1_km1 = cur.s1;
        double p_km2 = prim[q0 - sz], p_km1 = cur.pc, p_k = cur.pkp1;
        double q1_km1 = cur.q1, q2_km1 = cur.q2;
        double Om_k = w_above(cur, 1) - cur.szz;
        // ring of FD levels of loads in flight (slots indexed statically after unrolling)
        constexpr int FD = decltype(depth)::value;
        Fw ring[FD];
#pragma unroll
        for (int d = 0; d < FD; ++d)
            if (1 + d <= nz - 1) fload(q0 + (1 + d) * sz, ring[d]);
        for (int kb = 1; kb <= nz - 1; kb += FD) {
#pragma unroll
            for (int d = 0; d < FD; ++d) {
                const int k = kb + d;
                if (k > nz - 1) break;
                const long q = q0 + k * sz;
                Fw c = ring[d];
                if (k + FD <= nz - 1) fload(q + FD * sz, ring[d]);
                double t0_k, t1_k;
                hdiv(c, t0_k, t1_k);
                tr0[q] = t0_k;
                tr1[q] = t1_k;
                drw[q] = Om_k;
                double Om_kp1 = w_above(c, k + 1) - c.szz;
                double drho_k = c.pr - c.st;
                double drt_k = c.prt_ - c.srt_;
                double coeff_P = c.P, coeff_Q = c.Q;
                if (moist) {   // :379-384
                    double qf = 0.5 * (c.q1 + q1_km1 + c.q2 + q2_km1);
                    coeff_P /= (1.0 + qf);
                    coeff_Q /= (1.0 + qf);
                }
                double theta_t_lo = 0.5 * (p_km2 + p_km1);
                double theta_t_mid = 0.5 * (p_km1 + p_k);
                double theta_t_hi = 0.5 * (p_k + c.pkp1);
                double R0_tmp = coeff_P * drt_k + coeff_Q * drt_km1 - halfg * (drho_k + drho_km1);
                double R1_tmp = halfg * (-c.s0 - s0_km1 + t0_k + t0_km1) +
                                (coeff_P * (c.s1 - t1_k) + coeff_Q * (s1_km1 - t1_km1));
                R1_tmp += beta_1 * dzi * ((Om_kp1 - Om_km1) * halfg
                                          - (Om_kp1 * theta_t_hi - Om_k * theta_t_mid) * coeff_P
                                          - (Om_k * theta_t_mid - Om_km1 * theta_t_lo) * coeff_Q);
                double RHS = Om_k + dtau * (c.ws + R0_tmp + dtau * beta_2 * R1_tmp);
                double sol = implicit ? (RHS - c.A * sol_prev) * c.iB : RHS;
                soln[q] = sol;
                sol_prev = sol;
                p_km2 = p_km1; p_km1 = p_k; p_k = c.pkp1;
                q1_km1 = c.q1; q2_km1 = c.q2;
                Om_km1 = Om_k; Om_k = Om_kp1;
                drho_km1 = drho_k; drt_km1 = drt_k;
                s0_km1 = c.s0; s1_km1 = c.s1;
                t0_km1 = t0_k; t1_km1 = t1_k;
            }
        }
        // top face (:423-432): Om_k now holds face nz
        const long qt = q0 + nz * sz;
        double rhs_t = Om_k + dtau * slow_rhs.zmom[qt];
        double sol_t = implicit ? (rhs_t - A[qt] * sol_prev) * iB[qt] : rhs_t;
        data.zmom[qt] = stage.zmom[qt] + sol_t;
        drw[qt] = Om_k;

        // backward sweep with the cell update (K5)
        struct Bw { double so, C, iB, dw, szz, t0, t1, pc, pm, s0, s1, c0, c1; };
        auto bload = [&](long q, Bw& v) {
            v.so = soln[q]; v.C = Cc[q]; v.iB = iB[q]; v.dw = drw[q]; v.szz = stage.zmom[q];
            v.t0 = tr0[q]; v.t1 = tr1[q]; v.pc = prim[q]; v.pm = prim[q - sz]; v.s0 = sr0[q]; v.s1 = sr1[q];
            v.c0 = step == 0 ? prev.cons[q] : cc0[q];
            v.c1 = step == 0 ? prev.cons[q + ne] : cc1[q];
        };
        double sol_next = sol_t;
        double drw_next = Om_k;
        double p_next = prim[qt];
        constexpr int BD = decltype(depth)::value;
        Bw bring[BD];
#pragma unroll
        for (int d = 0; d < BD; ++d)
            if (nz - 1 - d >= 0) bload(q0 + (nz - 1 - d) * sz, bring[d]);
        for (int kb = nz - 1; kb >= 0; kb -= BD) {
#pragma unroll
            for (int d = 0; d < BD; ++d) {
                const int k = kb - d;
                if (k < 0) break;
                const long q = q0 + k * sz;
                Bw bc = bring[d];
                if (k - BD >= 0) bload(q - BD * sz, bring[d]);
                double sol = implicit ? bc.so - (bc.C * bc.iB) * sol_next : bc.so;
                double drw_k = bc.dw;
                data.zmom[q] = bc.szz + sol;
                double zflux_lo = beta_2 * sol + beta_1 * drw_k;
                double zflux_hi = beta_2 * sol_next + beta_1 * drw_next;
                scratch.zmom[q] += facinv * zflux_lo / (1.0 * 1.0);
                if (k == nz - 1) scratch.zmom[q + sz] += facinv * zflux_hi / (1.0 * 1.0);
                double pcq = bc.pc;
                double t0 = bc.t0 + dzi * (zflux_hi - zflux_lo);
                double t1 = bc.t1 + 0.5 * dzi * (zflux_hi * (pcq + p_next) - zflux_lo * (pcq + bc.pm));
                double r_new = bc.c0 + dtau * (bc.s0 - t0);
                double rt_new = bc.c1 + dtau * (bc.s1 - t1);
                cc0[q] = r_new;
                cc1[q] = rt_new;
                if (gw) {
                    // ghost layer 1 as the skipped apply_bcs would leave it: the
                    // periodic images of this cell, and at k = 0 / nz-1 the
                    // foextrap rows below / above, images included
                    const int ia[2] = {i, i == 0 ? nx : (i == nx - 1 ? -1 : i)};
                    const int ja[2] = {j, j == 0 ? ny : (j == ny - 1 ? -1 : j)};
                    const int nk = (k == 0 ? 1 : 0) + (k == nz - 1 ? 1 : 0);
                    const int ka[3] = {k, k == 0 ? -1 : nz, nz};
                    for (int a = 0; a < 2; ++a) {
                        if (a == 1 && ia[1] == i) continue;
                        for (int b = 0; b < 2; ++b) {
                            if (b == 1 && ja[1] == j) continue;
                            for (int c3 = 0; c3 <= nk; ++c3) {
                                if (a == 0 && b == 0 && c3 == 0) continue;
                                long g = L.ix(ia[a], ja[b], ka[c3]);
                                cc0[g] = r_new;
                                cc1[g] = rt_new;
                            }
                        }
                    }
                }
                sol_next = sol;
                drw_next = drw_k;
                p_next = pcq;
            }
        }
    };
    // The same column work as a tile of 64 columns per work-group of 8 waves
    // (launch_tiles), for boxes of up to 96 levels:
    //  1. each wave takes a chunk of faces k = ka..kb of its 64 columns and forms,
    //     level by level as the column kernel does, the horizontal divergences of
    //     cell k and Omega of face k (kept in the lane's registers) and the
    //     tridiagonal right-hand side of face k, staged in LDS with the face's
    //     a_k and 1/b_k;
    //  2. wave 0 runs the elimination up the 64 columns in LDS; the waves then
    //     stage c_k / b_k and Omega in LDS in place of a_k and 1/b_k, and wave 0
    //     runs the substitution down the columns;
    //  3. each wave updates rho*w, the averaged rho*w, rho and rho*theta of its
    //     chunk's cells from the solution in LDS.
    // Every value is computed by the same expression as in the column kernel, so
    // the results are bit-identical; the sweep's intermediates never reach HBM,
    // the per-level loads of a column are spread over 8 waves, and the serial
    // sweeps read only LDS.
    const bool implicit_ = implicit;
    auto tile = [=] __device__(int i, int j, bool valid, int lane, int wave, double* lds, auto cmc, auto tcc) {
        constexpr int CM = decltype(cmc)::value;
        constexpr int TC = decltype(tcc)::value;   // columns per tile
        constexpr int NW = 512 / TC;               // chunks of levels
        const int ch = tile_chunk(nz, NW);
        const int ka = wave * ch;
        const int kb = (ka + ch < nz ? ka + ch : nz) - 1;
        const bool work = valid && ka <= kb;
        const long q0 = valid ? L.ix(i, j, 0) : 0;
        constexpr int R = tile_rows<CM, NW>() * TC;
        double* X = lds;            // right-hand side, then the solution, per face
        double* Y = lds + R;        // a_k, then c_k / b_k
        double* Z = lds + 2 * R;    // 1 / b_k, then Omega
        struct Fw {
            double P, Q, ws, pkp1, pr, st, prt_, srt_, s0, s1;
            double dxl, dxh, sxl, sxh, dyl, dyh, syl, syh, pc, pxm, pxp, pym, pyp;
            double q1, q2;
        };
        auto fload = [&](long q, Fw& v) {
            v.P = Pc[q]; v.Q = Qc[q]; v.ws = slow_rhs.zmom[q];
            v.pkp1 = prim[q + sz];
            v.pr = pr0[q];
            v.st = st0[q]; v.prt_ = prt[q]; v.srt_ = srt[q]; v.s0 = sr0[q]; v.s1 = sr1[q];
            v.dxl = data.xmom[q]; v.dxh = data.xmom[q + 1]; v.sxl = stage.xmom[q]; v.sxh = stage.xmom[q + 1];
            v.dyl = data.ymom[q]; v.dyh = data.ymom[q + sy]; v.syl = stage.ymom[q]; v.syh = stage.ymom[q + sy];
            v.pc = prim[q]; v.pxm = prim[q - 1]; v.pxp = prim[q + 1]; v.pym = prim[q - sy]; v.pyp = prim[q + sy];
            if (moist) { v.q1 = pq1[q]; v.q2 = pq2[q]; } else { v.q1 = 0.0; v.q2 = 0.0; }
        };
        auto hdiv = [&](const Fw& v, double& t0, double& t1) {
            double xflux_lo = v.dxl - v.sxl;
            double xflux_hi = v.dxh - v.sxh;
            double yflux_lo = v.dyl - v.syl;
            double yflux_hi = v.dyh - v.syh;
            double pc = v.pc;
            t0 = (xflux_hi - xflux_lo) * dxi + (yflux_hi - yflux_lo) * dyi;
            t1 = ((xflux_hi * (pc + v.pxp) - xflux_lo * (pc + v.pxm)) * dxi +
                  (yflux_hi * (pc + v.pyp) - yflux_lo * (pc + v.pym)) * dyi) * 0.5;
        };
        // Omega of the old increments on face kf: prev rho*w (after the round
        // trip in the substep chain) minus the stage's
        auto om_face = [&](int kf) -> double {
            const long qf = q0 + (long)kf * sz;
            double pzv;
            if (kf == 0) {
                pzv = rt ? ch_w_lo * 0.5 * (pr0[q0] + pr0[q0]) : prev.zmom[q0];
            } else if (!rt) {
                pzv = prev.zmom[qf];
            } else if (kf >= nz) {
                pzv = ch_w_hi * 0.5 * (pr0[qf - sz] + pr0[qf - sz]);
            } else {
                double d = pr0[qf] + pr0[qf - sz];
                double w = prev.zmom[qf] * 2.0 / d;
                pzv = w * 0.5 * d;
            }
            return pzv - stage.zmom[qf];
        };
        double t0r[CM], t1r[CM], omr[CM + 1];
        // ---- phase 1: right-hand sides of the chunk's faces ----
        if (work) {
            double t0_km1, t1_km1, drho_km1, drt_km1, s0_km1, s1_km1, p_km2, p_km1, p_k, q1_km1, q2_km1;
            double Om_km1, Om_k;
            {
                const int kc = ka == 0 ? 0 : ka - 1;      // the cell below the chunk's first face
                Fw c;
                fload(q0 + (long)kc * sz, c);
                hdiv(c, t0_km1, t1_km1);
                drho_km1 = c.pr - c.st;
                drt_km1 = c.prt_ - c.srt_;
                s0_km1 = c.s0; s1_km1 = c.s1;
                p_km2 = prim[q0 + (long)(kc - 1) * sz]; p_km1 = c.pc; p_k = c.pkp1;
                q1_km1 = c.q1; q2_km1 = c.q2;
                Om_km1 = om_face(kc);
                Om_k = om_face(kc + 1);
                if (ka == 0) {      // face 0: the bottom row (:423-432)
                    X[lane] = Om_km1 + dtau * c.ws;
                    Z[lane] = iB[q0];
                    t0r[0] = t0_km1; t1r[0] = t1_km1; omr[0] = Om_km1;
                }
            }
#pragma unroll
            for (int m = 0; m < CM; ++m) {
                const int k = ka + m;
                if (k == 0) continue;
                if (k > kb) break;
                const long q = q0 + (long)k * sz;
                Fw c;
                fload(q, c);
                const double a_k = A[q], b_k = iB[q];
                double t0_k, t1_k;
                hdiv(c, t0_k, t1_k);
                double Om_kp1 = om_face(k + 1);
                double drho_k = c.pr - c.st;
                double drt_k = c.prt_ - c.srt_;
                double coeff_P = c.P, coeff_Q = c.Q;
                if (moist) {   // :379-384
                    double qf = 0.5 * (c.q1 + q1_km1 + c.q2 + q2_km1);
                    coeff_P /= (1.0 + qf);
                    coeff_Q /= (1.0 + qf);
                }
                double theta_t_lo = 0.5 * (p_km2 + p_km1);
                double theta_t_mid = 0.5 * (p_km1 + p_k);
                double theta_t_hi = 0.5 * (p_k + c.pkp1);
                double R0_tmp = coeff_P * drt_k + coeff_Q * drt_km1 - halfg * (drho_k + drho_km1);
                double R1_tmp = halfg * (-c.s0 - s0_km1 + t0_k + t0_km1) +
                                (coeff_P * (c.s1 - t1_k) + coeff_Q * (s1_km1 - t1_km1));
                R1_tmp += beta_1 * dzi * ((Om_kp1 - Om_km1) * halfg
                                          - (Om_kp1 * theta_t_hi - Om_k * theta_t_mid) * coeff_P
                                          - (Om_k * theta_t_mid - Om_km1 * theta_t_lo) * coeff_Q);
                X[k * TC + lane] = Om_k + dtau * (c.ws + R0_tmp + dtau * beta_2 * R1_tmp);
                Y[k * TC + lane] = a_k;
                Z[k * TC + lane] = b_k;
                t0r[m] = t0_k; t1r[m] = t1_k; omr[m] = Om_k;
                tile_level_fence();
                p_km2 = p_km1; p_km1 = p_k; p_k = c.pkp1;
                q1_km1 = c.q1; q2_km1 = c.q2;
                Om_km1 = Om_k; Om_k = Om_kp1;
                drho_km1 = drho_k; drt_km1 = drt_k;
                s0_km1 = c.s0; s1_km1 = c.s1;
                t0_km1 = t0_k; t1_km1 = t1_k;
            }
#pragma unroll
            for (int m = 0; m <= CM; ++m)      // Omega of the face above the chunk
                if (m == kb - ka + 1) omr[m] = Om_k;
            if (kb == nz - 1) {                // the top face's row (:423-432)
                const long qt = q0 + (long)nz * sz;
                X[nz * TC + lane] = Om_k + dtau * slow_rhs.zmom[qt];
                Y[nz * TC + lane] = A[qt];
                Z[nz * TC + lane] = iB[qt];
            }
        }
        __syncthreads();
        // ---- phase 2: elimination (wave 0), c_k / b_k and Omega staged, substitution (wave 0) ----
        if (wave == 0 && valid && implicit_) tile_forward<TC>(X, Y, Z, nz, lane);
        __syncthreads();
        if (wave == 0 && valid) {
            const long qt = q0 + (long)nz * sz;
            data.zmom[qt] = stage.zmom[qt] + X[nz * TC + lane];
        }
        if (work) {
#pragma unroll
            for (int m = 0; m < CM; ++m) {
                const int k = ka + m;
                if (k > kb) break;
                if (implicit_) Y[k * TC + lane] = Cc[q0 + (long)k * sz] * Z[k * TC + lane];
                Z[k * TC + lane] = omr[m];
            }
            if (kb == nz - 1) {
#pragma unroll
                for (int m = 0; m <= CM; ++m)
                    if (m == kb - ka + 1) Z[nz * TC + lane] = omr[m];
            }
        }
        __syncthreads();
        if (wave == 0 && valid && implicit_) tile_backward<TC>(X, Y, nz, lane);
        __syncthreads();
        // ---- phase 3: the chunk's cells ----
        if (work) {
#pragma unroll
            for (int m = 0; m < CM; ++m) {
                const int k = ka + m;
                if (k > kb) break;
                const long q = q0 + (long)k * sz;
                const double sol = X[k * TC + lane], sol_next = X[(k + 1) * TC + lane];
                const double szz = stage.zmom[q], azm = scratch.zmom[q];
                const double pcq = prim[q], p_next = prim[q + sz], pm = prim[q - sz];
                const double s0 = sr0[q], s1 = sr1[q];
                const double c0 = step == 0 ? prev.cons[q] : cc0[q];
                const double c1 = step == 0 ? prev.cons[q + ne] : cc1[q];
                double drw_k = Z[k * TC + lane], drw_next = Z[(k + 1) * TC + lane];
                data.zmom[q] = szz + sol;
                double zflux_lo = beta_2 * sol + beta_1 * drw_k;
                double zflux_hi = beta_2 * sol_next + beta_1 * drw_next;
                scratch.zmom[q] = azm + facinv * zflux_lo / (1.0 * 1.0);
                if (k == nz - 1) scratch.zmom[q + sz] += facinv * zflux_hi / (1.0 * 1.0);
                double t0 = t0r[m] + dzi * (zflux_hi - zflux_lo);
                double t1 = t1r[m] + 0.5 * dzi * (zflux_hi * (pcq + p_next) - zflux_lo * (pcq + pm));
                double r_new = c0 + dtau * (s0 - t0);
                double rt_new = c1 + dtau * (s1 - t1);
                cc0[q] = r_new;
                cc1[q] = rt_new;
                if (gw) {
                    const int ia[2] = {i, i == 0 ? nx : (i == nx - 1 ? -1 : i)};
                    const int ja[2] = {j, j == 0 ? ny : (j == ny - 1 ? -1 : j)};
                    const int nk = (k == 0 ? 1 : 0) + (k == nz - 1 ? 1 : 0);
                    const int ka3[3] = {k, k == 0 ? -1 : nz, nz};
                    for (int a = 0; a < 2; ++a) {
                        if (a == 1 && ia[1] == i) continue;
                        for (int b = 0; b < 2; ++b) {
                            if (b == 1 && ja[1] == j) continue;
                            for (int c3 = 0; c3 <= nk; ++c3) {
                                if (a == 0 && b == 0 && c3 == 0) continue;
                                long g = L.ix(ia[a], ja[b], ka3[c3]);
                                cc0[g] = r_new;
                                cc1[g] = rt_new;
                            }
                        }
                    }
                }
                tile_level_fence();
            }
        }
    };
    // the variant: the sim's choice, or the next one to time while tuning
    int var = P.col_variant;
    int tag = -1;
    if (P.tuner && P.tuner->chosen < 0 && P.tuner->next < COL_NV * ColTuner::SAMPLES) {
        tag = P.tuner->next++;
        var = tag % COL_NV;
        (void)hipEventRecord(P.tuner->ev[2 * tag], s);
    }
    // the tile form needs nz <= 96; a recorded (batched) launch takes the
    // largest chunk, so every member of a batch records the same kernel
    const int cm = tile_cm8(nz) == 0 ? 0 : (g_rec ? 12 : tile_cm8(nz));
    // 97..128 levels on a box of few columns: the tile form in 32- or
    // 16-column tiles (three arrays of 129 rows x 32 columns are 99 KB of LDS;
    // 64 columns would not fit).  A box of many columns keeps the per-lane
    // form: on the 256 x 256 x 128 CBL the autotuner prefers it to 32-column
    // tiles.
    const bool deep = cm == 0 && !g_rec && nz <= 128 && tile_columns((long)nx * ny) < 64;
    if (var == COL_TILE && cm == 0 && !deep) var = COL_FALLBACK;
    kt_start(P.timer, KT_FAST_COLUMN, s);
    if (var == COL_TILE) {
        auto launch = [&](auto cmc) {
            launch_tiles<decltype(cmc)::value, 8>(0, nx - 1, 0, ny - 1, s,
                [=] __device__(int i, int j, bool valid, int lane, int wave, double* lds) {
                    tile(i, j, valid, lane, wave, lds, cmc, std::integral_constant<int, 64>{});
                });
        };
        auto launch_n = [&](auto cmc, auto tcc) {
            launch_tiles_n<decltype(cmc)::value, decltype(tcc)::value>(0, nx - 1, 0, ny - 1, s,
                [=] __device__(int i, int j, bool valid, int lane, int wave, double* lds) {
                    tile(i, j, valid, lane, wave, lds, cmc, tcc);
                });
        };
        using IC4 = std::integral_constant<int, 4>;
        using IC8 = std::integral_constant<int, 8>;
        // a box of few columns or of more than 96 levels (a recorded launch
        // excepted) takes narrow tiles
        int tc = g_rec ? 64 : tile_columns((long)nx * ny);
        if (deep && tc > 32) tc = 32;
        if (tc == 16) {   // 32 chunks: at most 4 levels each up to 128 levels
            launch_n(IC4{}, std::integral_constant<int, 16>{});
        } else if (tc == 32) {   // 16 chunks: at most 8 levels each up to 128 levels
            if (tile_chunk(nz, 16) <= 4) launch_n(IC4{}, std::integral_constant<int, 32>{});
            else launch_n(IC8{}, std::integral_constant<int, 32>{});
        } else {
            switch (cm) {
                case 4: launch(IC4{}); break;
                case 8: launch(IC8{}); break;
                default: launch(std::integral_constant<int, 12>{}); break;
            }
        }
    } else {
        const int width = COL_WIDTH[var];
        switch (COL_DEPTH[var]) {
        case 2:
            launch_cols(0, nx - 1, 0, ny - 1, s,
                        [=] __device__(int i, int j) { column(i, j, std::integral_constant<int, 2>{}); }, width);
            break;
        case 4:
            launch_cols(0, nx - 1, 0, ny - 1, s,
                        [=] __device__(int i, int j) { column(i, j, std::integral_constant<int, 4>{}); }, width);
            break;
        default:
            launch_cols(0, nx - 1, 0, ny - 1, s,
                        [=] __device__(int i, int j) { column(i, j, std::integral_constant<int, 3>{}); }, width);
            break;
        }
    }
    if (tag >= 0) (void)hipEventRecord(P.tuner->ev[2 * tag + 1], s);
    // algorithmic bytes per cell: the fields of the substep this kernel reads
    // (stage rho, rho*theta, rho*u, rho*v, rho*w; the previous rho*w; the new
    // rho*u, rho*v; the current rho, rho*theta; the slow RHS of rho, rho*theta,
    // rho*w; the averaged rho*w; qv, qc when moist) and writes (rho*w, averaged
    // rho*w, rho, rho*theta), 8 B each.  The tridiagonal coefficients, theta
    // and the sweep's intermediates are not counted.
    kt_stop(P.timer, KT_FAST_COLUMN, s, (moist ? 20.0 : 18.0) * 8.0 * nx * ny * nz, "fast_rhs_N column");
}

}  // namespace erfx

// ---- erf_fast_rhs_N with no device allocation ---------------------------------
// For callers that cannot allocate device memory (code under stream / graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the work
// fields of the substep (theta extrapolated, the rho*w increment, the two
// horizontal divergences and the tridiagonal solution: 5 fields of the
// layout) live in one static device array instead of a context's workspace.
// Each GPU holds its own copy of the array until the process ends; calls on
// one GPU share it, so two calls that overlap on one GPU race.  No element of
// it needs to start at zero: every launch of the substep writes each element
// it reads (K1 the grown box of theta, the column kernel its own column)
// before reading it, so whatever a previous call left there is never seen.
namespace erfx {
namespace noalloc {
constexpr size_t SCRATCH_DOUBLES = size_t(8) << 20;      // 64 MiB
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];
// The array's device address per device, looked up once on the device's first
// call (the host-side name of a __device__ array is only a stand-in: passing it
// to a kernel would fault).
static double* g_ptr[MAX_DEVICES];
static std::mutex g_mu;

static int device_scratch(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("erfx_fast_rhs_N_noalloc: device index beyond the table of scratch addresses");
        return ERFX_ERR_NODEVICE;
    }
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return ERFX_OK;
}
}  // namespace noalloc

}  // namespace erfx

using namespace erfx;

extern "C" int erfx_fast_rhs_N_noalloc(const erfx_config* cfg, void* stream, int step, int nrk,
                                       const erfx_state* slow_rhs, const erfx_state* prev,
                                       const erfx_state* stage, const double* prim, const double* pi_stage,
                                       const double* fast_coeffs, erfx_state* data, erfx_state* scratch,
                                       double dtau, double beta_s, double facinv) {
    auto bad = [](const erfx_state* s) { return !s || !s->cons || !s->xmom || !s->ymom || !s->zmom; };
    if (!cfg || bad(slow_rhs) || bad(prev) || bad(stage) || bad(data) || bad(scratch) || !prim || !pi_stage ||
        !fast_coeffs) {
        set_error("null argument");
        return ERFX_ERR_ARG;
    }
    int rc = validate_config(*cfg);
    if (rc) return rc;
    if (cfg->use_terrain) { set_error("a terrain-following grid takes erf_fast_rhs_T"); return ERFX_ERR_STATE; }
    if (!(nrk > 0 || step == 0)) {
        set_error("the first RK stage takes a single substep (ERF_fast_rhs_N.cpp:106)");
        return ERFX_ERR_ARG;
    }
    erfx_layout lay;
    erfx_layout_for(cfg, &lay);
    const size_t ne = (size_t)lay.nelem;
    if (5 * ne > noalloc::SCRATCH_DOUBLES) {
        set_error("erfx_fast_rhs_N_noalloc: the box needs " + std::to_string(5 * ne) +
                  " doubles of work space, the static buffer holds " + std::to_string(noalloc::SCRATCH_DOUBLES) +
                  " (5 fields of the layout: boxes up to about 96 x 96 x 96 cells)");
        return ERFX_ERR_ARG;
    }
    double* base = nullptr;
    rc = noalloc::device_scratch(&base);
    if (rc) return rc;
    Params P;
    make_params(*cfg, lay, P);
    FastWork W;
    W.extrap = base;
    W.drw = base + ne;
    W.trhs = base + 2 * ne;      // two components
    W.soln = base + 4 * ne;
    fast_rhs_N(P, W, step, nrk, *slow_rhs, *prev, *stage, prim, pi_stage, fast_coeffs, *data, *scratch, dtau,
               beta_s, facinv, static_cast<hipStream_t>(stream));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    return ERFX_OK;
}
