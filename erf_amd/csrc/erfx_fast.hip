// erfx_fast.hip -- acoustic substep on a flat grid: make_fast_coeffs and
// erf_fast_rhs_N as gfx950 kernels.
//
// Launch structure per substep:
//   K1  cells grown by 1 + z-faces: Delta(rho theta), extrapolated theta,
//       lagged update, Delta(rho w) of the previous substep
//   K2  x- and y-faces (one launch): horizontal momentum update, written
//       straight into the current state (each lane reads prev and writes cur
//       at its own index, so the S_prev == S_data aliasing is safe)
//   K3  cells: horizontal flux divergence of the increments
//   K4  columns (one lane per (i, j), k loop): tridiagonal RHS, Thomas sweep,
//       new rho*w
//   K5  cells: vertical fluxes, time-averaged rho*w, rho and rho*theta update
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
        if (k >= 0) drw[q] = prev.zmom[q] - stage.zmom[q];
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

    // K3 (:328-357)
    launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        double xflux_lo = data.xmom[q] - stage.xmom[q];
        double xflux_hi = data.xmom[q + 1] - stage.xmom[q + 1];
        double yflux_lo = data.ymom[q] - stage.ymom[q];
        double yflux_hi = data.ymom[q + sy] - stage.ymom[q + sy];
        double pc = prim[q];
        tr0[q] = (xflux_hi - xflux_lo) * dxi + (yflux_hi - yflux_lo) * dyi;
        tr1[q] = ((xflux_hi * (pc + prim[q + 1]) - xflux_lo * (pc + prim[q - 1])) * dxi +
                  (yflux_hi * (pc + prim[q + sy]) - yflux_lo * (pc + prim[q - sy])) * dyi) * 0.5;
    });

    // K4 (:374-453): RHS on interior faces, boundary rows, Thomas sweep (GPU branch)
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
    kt_start(P.timer, KT_FAST_COLUMN, s);
    launch_cols(0, nx - 1, 0, ny - 1, s, [=] __device__(int i, int j) {
        const long q0 = L.ix(i, j, 0);
        // bottom row
        double rhs0 = prev.zmom[q0] - stage.zmom[q0] + dtau * slow_rhs.zmom[q0];
        // every RHS reads prev.zmom (== data.zmom for step > 0), so no rho*w is
        // written before the forward sweep is complete
        double sol_prev = rhs0 * iB[q0];
        soln[q0] = sol_prev;
        const double w0 = stage.zmom[q0] + sol_prev;
        for (int k = 1; k <= nz; ++k) {
            const long q = q0 + k * sz;
            double RHS;
            if (k <= nz - 1) {
                double coeff_P = Pc[q];
                double coeff_Q = Qc[q];
                double theta_t_lo = 0.5 * (prim[q - 2 * sz] + prim[q - sz]);
                double theta_t_mid = 0.5 * (prim[q - sz] + prim[q]);
                double theta_t_hi = 0.5 * (prim[q] + prim[q + sz]);
                double Omega_kp1 = prev.zmom[q + sz] - stage.zmom[q + sz];
                double Omega_k = prev.zmom[q] - stage.zmom[q];
                double Omega_km1 = prev.zmom[q - sz] - stage.zmom[q - sz];
                double old_drho_k = pr0[q] - st0[q];
                double old_drho_km1 = pr0[q - sz] - st0[q - sz];
                double drt_k = prt[q] - srt[q];
                double drt_km1 = prt[q - sz] - srt[q - sz];
                double R0_tmp = coeff_P * drt_k + coeff_Q * drt_km1 - halfg * (old_drho_k + old_drho_km1);
                double R1_tmp = halfg * (-sr0[q] - sr0[q - sz] + tr0[q] + tr0[q - sz]) +
                                (coeff_P * (sr1[q] - tr1[q]) + coeff_Q * (sr1[q - sz] - tr1[q - sz]));
                R1_tmp += beta_1 * dzi * ((Omega_kp1 - Omega_km1) * halfg
                                          - (Omega_kp1 * theta_t_hi - Omega_k * theta_t_mid) * coeff_P
                                          - (Omega_k * theta_t_mid - Omega_km1 * theta_t_lo) * coeff_Q);
                RHS = Omega_k + dtau * (slow_rhs.zmom[q] + R0_tmp + dtau * beta_2 * R1_tmp);
            } else {
                RHS = prev.zmom[q] - stage.zmom[q] + dtau * slow_rhs.zmom[q];
            }
            double sol;
            if (implicit) sol = (RHS - A[q] * sol_prev) * iB[q];
            else sol = RHS;
            soln[q] = sol;
            sol_prev = sol;
        }
        data.zmom[q0] = w0;
        if (implicit) {
            const long qt = q0 + nz * sz;
            data.zmom[qt] = stage.zmom[qt] + sol_prev;
            double sol_next = sol_prev;
            for (int k = nz - 1; k >= 0; --k) {
                const long q = q0 + k * sz;
                double sol = soln[q];
                sol -= (Cc[q] * iB[q]) * sol_next;
                soln[q] = sol;
                data.zmom[q] = stage.zmom[q] + sol;
                sol_next = sol;
            }
        } else {
            soln[q0] = rhs0;
            for (int k = 0; k <= nz; ++k) {
                const long q = q0 + k * sz;
                data.zmom[q] = stage.zmom[q] + soln[q];
            }
        }
    });
    // algorithmic traffic per z-face: 17 fields read once (A, 1/B, C, P, Q, prim,
    // prev/stage zmom, prev/stage rho and rho*theta, slow rho, rho*theta and zmom,
    // the two horizontal divergences), soln and the new zmom written once
    kt_stop(P.timer, KT_FAST_COLUMN, s, 19.0 * 8.0 * nx * ny * (nz + 1), "fast_rhs_N column solve");

    // K5 (:518-536 and :566-592)
    double* cc0 = data.cons;
    double* cc1 = data.cons + ne;
    const double* pc0 = prev.cons;
    const double* pc1 = prev.cons + ne;
    launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        double zflux_lo = beta_2 * soln[q] + beta_1 * drw[q];
        double zflux_hi = beta_2 * soln[q + sz] + beta_1 * drw[q + sz];
        scratch.zmom[q] += facinv * zflux_lo / (1.0 * 1.0);
        if (k == nz - 1) scratch.zmom[q + sz] += facinv * zflux_hi / (1.0 * 1.0);
        double pcq = prim[q];
        double t0 = tr0[q] + dzi * (zflux_hi - zflux_lo);
        double t1 = tr1[q] + 0.5 * dzi * (zflux_hi * (pcq + prim[q + sz]) - zflux_lo * (pcq + prim[q - sz]));
        if (step == 0) {
            cc0[q] = pc0[q] + dtau * (sr0[q] - t0);
            cc1[q] = pc1[q] + dtau * (sr1[q] - t1);
        } else {
            cc0[q] += dtau * (sr0[q] - t0);
            cc1[q] += dtau * (sr1[q] - t1);
        }
    });
}

}  // namespace erfx
