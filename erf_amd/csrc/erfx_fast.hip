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
    launch_cols(0, nx - 1, 0, ny - 1, s, [=] __device__(int i, int j) {
        const long q0 = L.ix(i, j, 0);
        // interior faces k = 1..nz-1
        for (int k = 1; k <= nz - 1; ++k) {
            const long q = q0 + k * sz;
            double rhobar_lo = r0[q - sz], rhobar_hi = r0[q];
            double pibar_lo = pi0[q - sz], pibar_hi = pi0[q];
            double pi_c = 0.5 * (pi_stage[q - sz] + pi_stage[q]);
            double coeff_P = -Gamma * R_d * dzi + halfg * R_d * rhobar_hi / (c_v * pibar_hi * rt[q]);
            coeff_P *= pi_c;
            double coeff_Q = Gamma * R_d * dzi + halfg * R_d * rhobar_lo / (c_v * pibar_lo * rt[q - sz]);
            coeff_Q *= pi_c;
            Pc[q] = coeff_P;
            Qc[q] = coeff_Q;
            double theta_t_lo = 0.5 * (prim[q - 2 * sz] + prim[q - sz]);
            double theta_t_mid = 0.5 * (prim[q - sz] + prim[q]);
            double theta_t_hi = 0.5 * (prim[q] + prim[q + sz]);
            A[q] = D * (halfg - coeff_Q * theta_t_lo);
            Cc[q] = D * (-halfg + coeff_P * theta_t_hi);
            B[q] = 1.0 + D * (coeff_Q - coeff_P) * theta_t_mid;
        }
        A[q0] = 0.0; B[q0] = 1.0; Cc[q0] = 0.0;
        const long qt = q0 + nz * sz;
        A[qt] = top_outflow ? -1.0 : 0.0; B[qt] = 1.0; Cc[qt] = 0.0;
        double bet = B[q0];
        double c_prev = Cc[q0];
        for (int k = 1; k <= nz; ++k) {
            const long q = q0 + k * sz;
            double gam = c_prev / bet;
            bet = B[q] - A[q] * gam;
            c_prev = Cc[q];
            // store 1/B on interior faces only; the top keeps B itself (:287-290)
            B[q] = (k <= nz - 1) ? 1.0 / bet : bet;
        }
    });
}

void fast_rhs_N(const Params& P, const FastWork& W, int step, int nrk, const erfx_state& slow_rhs,
                const erfx_state& prev, const erfx_state& stage, const double* prim, const double* pi_stage,
                const double* fc, const erfx_state& data, const erfx_state& scratch, double dtau,
                double beta_s, double facinv, hipStream_t s) {
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

    // K2 (:161-257)
    const bool first = (nrk == 0 && step == 0);
    const double gr = -Gamma * R_d;
    launch_box(Box(0, 0, 0, nx, ny, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (j <= ny - 1) {
            double new_drho_u;
            if (first) {
                new_drho_u = prev.xmom[q] - stage.xmom[q] + dtau * slow_rhs.xmom[q];
            } else {
                double gpx = (extrap[q] - extrap[q - 1]) * dxi;
                double pi_c = 0.5 * (pi_stage[q - 1] + pi_stage[q]);
                double fast_rhs_rho_u = gr * pi_c * gpx;
                new_drho_u = prev.xmom[q] - stage.xmom[q] + dtau * fast_rhs_rho_u + dtau * slow_rhs.xmom[q];
            }
            scratch.xmom[q] += facinv * new_drho_u;
            data.xmom[q] = stage.xmom[q] + new_drho_u;
        }
        if (i <= nx - 1) {
            double new_drho_v;
            if (first) {
                new_drho_v = prev.ymom[q] - stage.ymom[q] + dtau * slow_rhs.ymom[q];
            } else {
                double gpy = (extrap[q] - extrap[q - sy]) * dyi;
                double pi_c = 0.5 * (pi_stage[q - sy] + pi_stage[q]);
                double fast_rhs_rho_v = gr * pi_c * gpy;
                new_drho_v = prev.ymom[q] - stage.ymom[q] + dtau * fast_rhs_rho_v + dtau * slow_rhs.ymom[q];
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
    kt_start(P.timer, KT_FAST_COLUMN, s);
    launch_cols(0, nx - 1, 0, ny - 1, s, [=] __device__(int i, int j) {
        const long q0 = L.ix(i, j, 0);
        // per-level loads of the forward sweep (face k and cell k)
        struct Fw {
            double P, Q, A, iB, ws, pkp1, pz, szz, pr, st, prt_, srt_, s0, s1;
            double dxl, dxh, sxl, sxh, dyl, dyh, syl, syh, pc, pxm, pxp, pym, pyp;
        };
        auto fload = [&](long q, Fw& v) {
            v.P = Pc[q]; v.Q = Qc[q]; v.A = A[q]; v.iB = iB[q]; v.ws = slow_rhs.zmom[q];
            v.pkp1 = prim[q + sz]; v.pz = prev.zmom[q + sz]; v.szz = stage.zmom[q + sz];
            v.pr = pr0[q]; v.st = st0[q]; v.prt_ = prt[q]; v.srt_ = srt[q]; v.s0 = sr0[q]; v.s1 = sr1[q];
            v.dxl = data.xmom[q]; v.dxh = data.xmom[q + 1]; v.sxl = stage.xmom[q]; v.sxh = stage.xmom[q + 1];
            v.dyl = data.ymom[q]; v.dyh = data.ymom[q + sy]; v.syl = stage.ymom[q]; v.syh = stage.ymom[q + sy];
            v.pc = prim[q]; v.pxm = prim[q - 1]; v.pxp = prim[q + 1]; v.pym = prim[q - sy]; v.pyp = prim[q + sy];
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

        Fw cur, nxt;
        fload(q0, cur);
        // face 0 (bottom row, :423-432) and cell 0
        double Om_km1 = prev.zmom[q0] - stage.zmom[q0];
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
        double Om_k = cur.pz - cur.szz;
        if (nz > 1) fload(q0 + sz, nxt);
        for (int k = 1; k <= nz - 1; ++k) {
            const long q = q0 + k * sz;
            cur = nxt;
            if (k + 1 <= nz - 1) fload(q + sz, nxt);
            double t0_k, t1_k;
            hdiv(cur, t0_k, t1_k);
            tr0[q] = t0_k;
            tr1[q] = t1_k;
            double Om_kp1 = cur.pz - cur.szz;
            double drho_k = cur.pr - cur.st;
            double drt_k = cur.prt_ - cur.srt_;
            double coeff_P = cur.P, coeff_Q = cur.Q;
            double theta_t_lo = 0.5 * (p_km2 + p_km1);
            double theta_t_mid = 0.5 * (p_km1 + p_k);
            double theta_t_hi = 0.5 * (p_k + cur.pkp1);
            double R0_tmp = coeff_P * drt_k + coeff_Q * drt_km1 - halfg * (drho_k + drho_km1);
            double R1_tmp = halfg * (-cur.s0 - s0_km1 + t0_k + t0_km1) +
                            (coeff_P * (cur.s1 - t1_k) + coeff_Q * (s1_km1 - t1_km1));
            R1_tmp += beta_1 * dzi * ((Om_kp1 - Om_km1) * halfg
                                      - (Om_kp1 * theta_t_hi - Om_k * theta_t_mid) * coeff_P
                                      - (Om_k * theta_t_mid - Om_km1 * theta_t_lo) * coeff_Q);
            double RHS = Om_k + dtau * (cur.ws + R0_tmp + dtau * beta_2 * R1_tmp);
            double sol = implicit ? (RHS - cur.A * sol_prev) * cur.iB : RHS;
            soln[q] = sol;
            sol_prev = sol;
            p_km2 = p_km1; p_km1 = p_k; p_k = cur.pkp1;
            Om_km1 = Om_k; Om_k = Om_kp1;
            drho_km1 = drho_k; drt_km1 = drt_k;
            s0_km1 = cur.s0; s1_km1 = cur.s1;
            t0_km1 = t0_k; t1_km1 = t1_k;
        }
        // top face (:423-432): Om_k now holds face nz
        const long qt = q0 + nz * sz;
        double rhs_t = Om_k + dtau * slow_rhs.zmom[qt];
        double sol_t = implicit ? (rhs_t - A[qt] * sol_prev) * iB[qt] : rhs_t;
        data.zmom[qt] = stage.zmom[qt] + sol_t;

        // backward sweep with the cell update (K5)
        struct Bw { double so, C, iB, pz, szz, t0, t1, pc, pm, s0, s1, c0, c1; };
        auto bload = [&](long q, Bw& v) {
            v.so = soln[q]; v.C = Cc[q]; v.iB = iB[q]; v.pz = prev.zmom[q]; v.szz = stage.zmom[q];
            v.t0 = tr0[q]; v.t1 = tr1[q]; v.pc = prim[q]; v.pm = prim[q - sz]; v.s0 = sr0[q]; v.s1 = sr1[q];
            v.c0 = step == 0 ? prev.cons[q] : cc0[q];
            v.c1 = step == 0 ? prev.cons[q + ne] : cc1[q];
        };
        double sol_next = sol_t;
        double drw_next = Om_k;
        double p_next = prim[qt];
        Bw bc, bn;
        bload(qt - sz, bn);
        for (int k = nz - 1; k >= 0; --k) {
            const long q = q0 + k * sz;
            bc = bn;
            if (k >= 1) bload(q - sz, bn);
            double sol = implicit ? bc.so - (bc.C * bc.iB) * sol_next : bc.so;
            double drw_k = bc.pz - bc.szz;
            data.zmom[q] = bc.szz + sol;
            double zflux_lo = beta_2 * sol + beta_1 * drw_k;
            double zflux_hi = beta_2 * sol_next + beta_1 * drw_next;
            scratch.zmom[q] += facinv * zflux_lo / (1.0 * 1.0);
            if (k == nz - 1) scratch.zmom[q + sz] += facinv * zflux_hi / (1.0 * 1.0);
            double pcq = bc.pc;
            double t0 = bc.t0 + dzi * (zflux_hi - zflux_lo);
            double t1 = bc.t1 + 0.5 * dzi * (zflux_hi * (pcq + p_next) - zflux_lo * (pcq + bc.pm));
            cc0[q] = bc.c0 + dtau * (bc.s0 - t0);
            cc1[q] = bc.c1 + dtau * (bc.s1 - t1);
            sol_next = sol;
            drw_next = drw_k;
            p_next = pcq;
        }
    });
    // algorithmic traffic per z-face: 22 fields read once (A, 1/B, C, P, Q, theta,
    // prev/stage rho*w, prev/stage rho and rho*theta, slow rho, rho*theta and
    // rho*w, new and stage x- and y-momenta, rho and rho*theta of the state being
    // updated, averaged rho*w) and 4 written (rho*w, averaged rho*w, rho, rho*theta)
    kt_stop(P.timer, KT_FAST_COLUMN, s, 26.0 * 8.0 * nx * ny * (nz + 1), "fast_rhs_N column (K3+K4+K5)");
}

}  // namespace erfx
