// erfx_terrain.hip -- terrain-following grid: the metrics of the node heights
// and the acoustic substep erf_fast_rhs_T (static terrain, unit map factors).
//
// The substep runs as three phases: theta extrapolation and the horizontal
// momentum increments on the grown box; the new horizontal momenta; and one
// lane per column that forms Omega of the old increments and the horizontal
// divergence of the new ones as it sweeps up, solves the implicit vertical
// system and updates rho, rho*theta and rho*w on the way down.  Every
// expression keeps the reference's operation order (FP contraction is off),
// so each point carries the reference's rounding.
#include "erfx_internal.h"
#include "erfx_terrain.h"

namespace erfx {

// one lane per element of the layout, padding included
template <class F>
__global__ void __launch_bounds__(256) lin_kernel(long n, F g) {
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q < n) g(q);
}

// make_J, make_areas, make_zcc (Source/Utils/ERF_TerrainMetrics.cpp:583-715):
// detJ and z_cc on cells (-1..nx, -1..ny, 0..nz), ax on x-faces (-1..nx+1),
// ay on y-faces (-1..ny+1); 1 (0 for z_cc) everywhere else, az = 1.
void terrain_metrics(const Params& P, const double* z_nd, double* detJ, double* ax, double* ay, double* az,
                     double* z_cc, double* mrx, double* mry, hipStream_t s) {
    if (rec_refuse("terrain metrics")) return;
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz, ng = L.ng;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const double dzInv = 1.0 / P.cfg.dz;
    const double dxi = P.dxInv[0], dyi = P.dxInv[1], dzi = P.dxInv[2];
    const Zn z{z_nd, L};
    const unsigned nb = (unsigned)((ne + 255) / 256);
    auto f = [=] __device__(long q) {
        const int k = (int)(q / sz) - ng;
        const long rem = q % sz;
        const int j = (int)(rem / sy) - ng, i = (int)(rem % sy) - ng;
        const bool kin = k >= 0 && k <= nz;
        double J = 1.0, Zc = 0.0, Ax = 1.0, Ay = 1.0;
        if (kin && i >= -1 && i <= nx && j >= -1 && j <= ny) {
            J = .25 * dzInv * (z(i, j, k + 1) + z(i + 1, j, k + 1) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1)
                               - z(i, j, k) - z(i + 1, j, k) - z(i, j + 1, k) - z(i + 1, j + 1, k));
            Zc = .125 * (z(i, j, k) + z(i + 1, j, k) + z(i, j + 1, k) + z(i + 1, j + 1, k)
                         + z(i, j, k + 1) + z(i + 1, j, k + 1) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1));
        }
        if (kin && i >= -1 && i <= nx + 1 && j >= -1 && j <= ny)
            Ax = .5 * dzInv * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k) - z(i, j + 1, k));
        if (kin && i >= -1 && i <= nx && j >= -1 && j <= ny + 1)
            Ay = .5 * dzInv * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k) - z(i + 1, j, k));
        detJ[q] = J; z_cc[q] = Zc; ax[q] = Ax; ay[q] = Ay; az[q] = 1.0;
        const bool kv = k >= 0 && k <= nz - 1;
        mrx[q] = (kv && i >= 0 && i <= nx && j >= 0 && j <= ny - 1)
                     ? met::h_xi_Iface(z, i, j, k, dxi) / met::h_zeta_Iface(z, i, j, k, dzi) : 0.0;
        mry[q] = (kv && i >= 0 && i <= nx - 1 && j >= 0 && j <= ny)
                     ? met::h_eta_Jface(z, i, j, k, dyi) / met::h_zeta_Jface(z, i, j, k, dzi) : 0.0;
    };
    hipLaunchKernelGGL(lin_kernel<decltype(f)>, dim3(nb), dim3(256), 0, s, ne, f);
}

// erf_fast_rhs_T (Source/TimeIntegration/ERF_fast_rhs_T.cpp:39-704), GPU
// branch of the column solve.  For step > 0 prev and data alias (S_sum).
void fast_rhs_T(const Params& P, const FastWork& W, const TerrainWork& T, int step, int nrk,
                const erfx_state& slow_rhs, const erfx_state& prev, const erfx_state& stage, const double* prim,
                const double* pi_stage, const double* fc, const erfx_state& data, const erfx_state& scratch,
                double dtau, double beta_s, double facinv, hipStream_t s, const FastChain& ch) {
    (void)nrk;
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const double beta_1 = 0.5 * (1.0 - beta_s);
    const double beta_2 = 0.5 * (1.0 + beta_s);
    const double beta_d = 0.1;
    const double dxi = P.dxInv[0], dyi = P.dxInv[1], dzi = P.dxInv[2];
    const bool moist = P.moist;
    const Zn z{P.z_nd, L};
    const double* J = P.detJ;
    double* cc0 = data.cons;
    double* cc1 = data.cons + ne;
    const double* pc0 = prev.cons;
    const double* pc1 = prev.cons + ne;
    const double* sc0 = stage.cons;
    const double* sc1 = stage.cons + ne;
    double* lagged1 = scratch.cons + ne;
    double* du = T.du; double* dv = T.dv;
    double* nu = T.nu; double* nv = T.nv; double* om = T.om;
    double* wtop = T.wtop;
    double* extrap = W.extrap;
    double* trhs0 = W.trhs;
    double* trhs1 = W.trhs + ne;
    double* soln = W.soln;
    const double* pxm = prev.xmom; const double* pym = prev.ymom; const double* pzm = prev.zmom;
    const double* sxm = stage.xmom; const double* sym = stage.ymom; const double* szm = stage.zmom;
    const bool first = step == 0;
    // The substep chain (laterally periodic box, Dirichlet top): prev is S_sum as the
    // previous substep left it, without the apply_bcs between the two.  What
    // that apply_bcs would give, the kernels form as they read: ghost cells of
    // rho*theta are the periodic images (foextrap rows above and below), and
    // each momentum is put through the momentum -> velocity -> momentum round
    // trip with the face density of its periodic image
    // (ERF_MomentumToVelocity.cpp, ERF_VelocityToMomentum.cpp), so the rounding
    // is the reference's.
    const bool rt = ch.rt_prev && step > 0;
    const double* prho = prev.cons;
    auto wrap_c = [=] __device__(int a, int n) { return a < 0 ? a + n : (a > n - 1 ? a - n : a); };
    auto wrap_f = [=] __device__(int a, int n) { return a < 0 ? a + n : (a > n ? a - n : a); };
    // the cell holding the value of (possibly ghost) cell (i, j, k)
    auto cimg = [=] __device__(int i, int j, int k) {
        return L.ix(wrap_c(i, nx), wrap_c(j, ny), k < 0 ? 0 : (k > nz - 1 ? nz - 1 : k));
    };
    // round trip of x-momentum on (possibly ghost) face (i, j, k), k valid
    auto rt_x = [=] __device__(int i, int j, int k) {
        const int fi = wrap_f(i, nx), cj = wrap_c(j, ny);
        const double d = prho[L.ix(wrap_c(fi, nx), cj, k)] + prho[L.ix(wrap_c(fi - 1, nx), cj, k)];
        const double u = pxm[L.ix(fi, cj, k)] * 2.0 / d;
        return u * 0.5 * d;
    };
    auto rt_y = [=] __device__(int i, int j, int k) {
        const int ci = wrap_c(i, nx), fj = wrap_f(j, ny);
        const double d = prho[L.ix(ci, wrap_c(fj, ny), k)] + prho[L.ix(ci, wrap_c(fj - 1, ny), k)];
        const double v = pym[L.ix(ci, fj, k)] * 2.0 / d;
        return v * 0.5 * d;
    };
    // rho*w on the top face of column (i, j) as the skipped apply_bcs leaves it
    // (Dirichlet top on a terrain grid, ERF_BoundaryConditions_zvel.cpp:220-257):
    // w = WFromOmega of the top value with u and v as the fill makes them --
    // prev's momenta over the face density of the images, the level above the
    // top from the xvel / yvel top rules -- then rho*w with the top cell's
    // density (the cell above it is a copy).  On a plane top the slopes vanish
    // and w is the Dirichlet value.
    const int xth = P.bck[0][5], yth = P.bck[1][5];
    const double xvh = P.vel_extdir[0][5], yvh = P.vel_extdir[1][5], zvh = P.vel_extdir[2][5];
    auto u_fill = [=] __device__(int ii, int jj, int kk) -> double {
        if (kk > nz - 1) {
            if (xth == EXT_DIR) return xvh;
            kk = nz - 1;
        }
        const int fi = wrap_f(ii, nx), cj = wrap_c(jj, ny);
        const double d = prho[L.ix(wrap_c(fi, nx), cj, kk)] + prho[L.ix(wrap_c(fi - 1, nx), cj, kk)];
        return pxm[L.ix(fi, cj, kk)] * 2.0 / d;
    };
    auto v_fill = [=] __device__(int ii, int jj, int kk) -> double {
        if (kk > nz - 1) {
            if (yth == EXT_DIR) return yvh;
            kk = nz - 1;
        }
        const int ci = wrap_c(ii, nx), fj = wrap_f(jj, ny);
        const double d = prho[L.ix(ci, wrap_c(fj, ny), kk)] + prho[L.ix(ci, wrap_c(fj - 1, ny), kk)];
        return pym[L.ix(ci, fj, kk)] * 2.0 / d;
    };
    auto pw_top = [=] __device__(int i, int j) -> double {
        const double w = met::w_from_omega(i, j, nz, zvh, u_fill, v_fill, z, dxi, dyi);
        const double r = prho[L.ix(i, j, nz - 1)];
        return w * 0.5 * (r + r);
    };

    // :139-204: rho, rho*theta copied from prev on the grown box (step 0);
    // theta extrapolated forward and the lagged increment; the horizontal
    // momentum increments on faces grown by one laterally, kept because the
    // new momenta overwrite prev when it aliases data (the rho, rho*theta and
    // rho*w increments are formed where they are used, from prev and stage)
    launch_box(Box(-1, -1, -1, nx + 1, ny, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (i <= nx) {
            if (first) { cc0[q] = pc0[q]; cc1[q] = pc1[q]; }
            double dt_ = (rt ? cc1[cimg(i, j, k)] : cc1[q]) - sc1[q];
            extrap[q] = first ? dt_ : dt_ + beta_d * (dt_ - lagged1[q]);
            lagged1[q] = dt_;
        }
        if (k >= 0) du[q] = (rt ? rt_x(i, j, k) : pxm[q]) - sxm[q];
        if (k >= 0 && i <= nx) dv[q] = (rt ? rt_y(i, j, k) : pym[q]) - sym[q];
    });
    // the top layer of cells (k = nz); in the chain also rho*w of the top face
    // of every column (read by the column kernel, after the horizontal
    // momenta and the densities of prev are overwritten)
    launch_box(Box(-1, -1, nz, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (first) { cc0[q] = pc0[q]; cc1[q] = pc1[q]; }
        double dt_ = (rt ? cc1[cimg(i, j, k)] : cc1[q]) - sc1[q];
        extrap[q] = first ? dt_ : dt_ + beta_d * (dt_ - lagged1[q]);
        lagged1[q] = dt_;
        if (rt && i >= 0 && i <= nx - 1 && j >= 0 && j <= ny - 1) wtop[L.ix(i, j, -L.ng)] = pw_top(i, j);
    });


    // :262-331 horizontal momenta with the terrain-corrected pressure gradient
    const double* pq1 = prim + PRIM_Q1 * ne;
    const double* pq2 = prim + PRIM_Q2 * ne;
    const double* srx = slow_rhs.xmom;
    const double* sry = slow_rhs.ymom;
    double* axm = scratch.xmom;
    double* aym = scratch.ymom;
    double* dxm = data.xmom;
    double* dym = data.ymom;
    const double* mrx = P.mrx;
    const double* mry = P.mry;
    // Each lane marches up a chunk of KC levels of its (i, j) and keeps extrap
    // of its own column and of the columns at i-1 and j-1 for levels k-1, k
    // and k+1 in registers, so every extrap value is read once per chunk
    // instead of at three levels (the box order would put the levels of the
    // gradient's stencil a whole plane of tiles apart).
    constexpr int KC = 24;
    const int nch = (nz + KC - 1) / KC;
    launch_box(Box(0, 0, 0, nx, ny, nch - 1), s, [=] __device__(int i, int j, int c) {
        const int k0 = c * KC, k1 = k0 + KC < nz ? k0 + KC : nz;
        const bool xf = j <= ny - 1, yf = i <= nx - 1;
        long q = L.ix(i, j, k0);
        // c: column (i, j); x: (i-1, j); y: (i, j-1); m / 0 / p: levels k-1 / k / k+1
        double ecm = 0.0, exm = 0.0, eym = 0.0;
        if (k0 > 0) {
            ecm = extrap[q - sz];
            if (xf) exm = extrap[q - 1 - sz];
            if (yf) eym = extrap[q - sy - sz];
        }
        double ec0 = extrap[q];
        double ex0 = xf ? extrap[q - 1] : 0.0;
        double ey0 = yf ? extrap[q - sy] : 0.0;
        for (int k = k0; k < k1; ++k, q += sz) {
            const double ecp = extrap[q + sz];
            const double exp_ = xf ? extrap[q - 1 + sz] : 0.0;
            const double eyp = yf ? extrap[q - sy + sz] : 0.0;
            if (xf) {
                double gp_xi = (ec0 - ex0) * dxi;
                double gp_zeta_on_iface = (k == 0)
                    ? 0.5 * dzi * (exp_ + ecp - ex0 - ec0)
                    : 0.25 * dzi * (exp_ + ecp - exm - ecm);
                double gpx = gp_xi - mrx[q] * gp_zeta_on_iface;   // (h_xi / h_zeta) made once
                gpx *= 1.0;
                if (moist) gpx /= (1.0 + face_q(pq1, pq2, q, q - 1));
                double pi_c = 0.5 * (pi_stage[q - 1] + pi_stage[q]);
                double fast_rhs_rho_u = -Gamma * R_d * pi_c * gpx;
                double n_ = du[q] + dtau * fast_rhs_rho_u + dtau * srx[q];
                nu[q] = n_;
                if (k == nz - 1) nu[q + sz] = n_;
                axm[q] += facinv * n_;
                dxm[q] = sxm[q] + n_;
            }
            if (yf) {
                double gp_eta = (ec0 - ey0) * dyi;
                double gp_zeta_on_jface = (k == 0)
                    ? 0.5 * dzi * (ecp + eyp - ec0 - ey0)
                    : 0.25 * dzi * (ecp + eyp - ecm - eym);
                double gpy = gp_eta - mry[q] * gp_zeta_on_jface;   // (h_eta / h_zeta) made once
                gpy *= 1.0;
                if (moist) gpy /= (1.0 + face_q(pq1, pq2, q, q - sy));
                double pi_c = 0.5 * (pi_stage[q - sy] + pi_stage[q]);
                double fast_rhs_rho_v = -Gamma * R_d * pi_c * gpy;
                double n_ = dv[q] + dtau * fast_rhs_rho_v + dtau * sry[q];
                nv[q] = n_;
                if (k == nz - 1) nv[q + sz] = n_;
                aym[q] += facinv * n_;
                dym[q] = sym[q] + n_;
            }
            ecm = ec0; ec0 = ecp;
            exm = ex0; ex0 = exp_;
            eym = ey0; ey0 = eyp;
        }
    });

    // :414-669, one lane per (i, j) column.  The forward sweep makes, level by
    // level, Omega of the old increments (:460-484) on the face above, the
    // horizontal divergence of the area-weighted new increments (:414-457) in
    // the cell, the tridiagonal right-hand side on the face (:499-552) and the
    // elimination (:561-581, GPU branch); Omega, the divergences and the
    // solution are kept for the backward sweep: substitution, rho*w = stage +
    // WFromOmega on the interior faces (faces 0 and nz keep the stage value,
    // :620-633), the time-averaged Omega and the rho, rho*theta update (:640-669).
    const Fld fdu{du, L}, fdv{dv, L}, fnu{nu, L}, fnv{nv, L};
    const double halfg = fabs(0.5 * (-P.gravity));
    const double* sr0 = slow_rhs.cons;
    const double* sr1 = slow_rhs.cons + ne;
    const double* srz = slow_rhs.zmom;
    const double* fA = fc;
    const double* fB = fc + ne;
    const double* fC = fc + 2 * ne;
    const double* fP = fc + 3 * ne;
    const double* fQ = fc + 4 * ne;
    double* cz = data.zmom;
    double* azm = scratch.zmom;
    kt_start(P.timer, KT_FAST_COLUMN, s);
    // The tile form (launch_tiles; the default for boxes of up to 96 levels):
    // the 8 waves of a 64-column tile form the right-hand sides of their
    // chunks of faces in parallel and stage them in LDS with a_k and 1/b_k
    // (Omega and the horizontal divergences stay in registers), wave 0 runs the
    // elimination in LDS, the waves stage c_k / b_k and Omega, wave 0 runs the
    // substitution and writes the top face, and the waves update their chunks'
    // cells.  Every value is the column kernel's expression, so the results are
    // bit-identical; the serial sweeps read only LDS.
    auto tile = [=] __device__(int i, int j, bool valid, int lane, int wave, double* lds, auto cmc) {
        constexpr int CM = decltype(cmc)::value;
        constexpr int NW = 8;
        const int ch = tile_chunk(nz, NW);
        const int ka = wave * ch;
        const int kb = (ka + ch < nz ? ka + ch : nz) - 1;
        const bool work = valid && ka <= kb;
        const long q0 = valid ? L.ix(i, j, 0) : 0;
        constexpr int R = tile_rows<CM, NW>() * 64;
        double* X = lds;            // right-hand side, then the solution, per face
        double* Y = lds + R;        // a_k, then c_k / b_k
        double* Z = lds + 2 * R;    // 1 / b_k, then Omega
        auto pw = [&](int kf) -> double {
            const long q = q0 + (long)kf * sz;
            if (!rt) return pzm[q];
            if (kf >= nz) return wtop[L.ix(i, j, -L.ng)];
            const double d = prho[q] + prho[q - sz];
            const double w = pzm[q] * 2.0 / d;
            return w * 0.5 * d;
        };
        auto omega_old = [&](int kf) -> double {
            const long q = q0 + (long)kf * sz;
            if (kf == 0) return 0.;
            const double dwk = pw(kf) - szm[q];
            if (kf == nz) return dwk;
            return met::omega_from_w(i, j, kf, dwk, fdu, fdv, z, dxi, dyi);
        };
        auto trhs_at = [&](int k, double& t0, double& t1) {
            const long q = q0 + (long)k * sz;
            double h_zeta_cc_xface_hi = 0.5 * dzi * (z(i + 1, j, k + 1) + z(i + 1, j + 1, k + 1) - z(i + 1, j, k) - z(i + 1, j + 1, k));
            double h_zeta_cc_xface_lo = 0.5 * dzi * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k) - z(i, j + 1, k));
            double h_zeta_cc_yface_hi = 0.5 * dzi * (z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1) - z(i, j + 1, k) - z(i + 1, j + 1, k));
            double h_zeta_cc_yface_lo = 0.5 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k) - z(i + 1, j, k));
            double xflux_lo = nu[q] * h_zeta_cc_xface_lo / 1.0;
            double xflux_hi = nu[q + 1] * h_zeta_cc_xface_hi / 1.0;
            double yflux_lo = nv[q] * h_zeta_cc_yface_lo / 1.0;
            double yflux_hi = nv[q + sy] * h_zeta_cc_yface_hi / 1.0;
            double mfsq = 1.0 * 1.0;
            t0 = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq;
            t1 = ((xflux_hi * (prim[q] + prim[q + 1]) - xflux_lo * (prim[q] + prim[q - 1])) * dxi * mfsq +
                  (yflux_hi * (prim[q] + prim[q + sy]) - yflux_lo * (prim[q] + prim[q - sy])) * dyi * mfsq) * 0.5;
        };
        double t0r[CM], t1r[CM], omr[CM + 1], pxr[CM] = {}, pyr[CM] = {};
        // ---- phase 1: right-hand sides of the chunk's faces ----
        if (work) {
            const int kc = ka == 0 ? 0 : ka - 1;      // the cell below the chunk's first face
            const long qc = q0 + (long)kc * sz;
            double t0_km1, t1_km1;
            trhs_at(kc, t0_km1, t1_km1);
            double drho_km1 = cc0[qc] - sc0[qc], drt_km1 = cc1[qc] - sc1[qc];
            double J_km1 = J[qc], s0_km1 = sr0[qc], s1_km1 = sr1[qc];
            double p_km2 = prim[qc - sz], p_km1 = prim[qc];
            double q1_km1 = moist ? pq1[qc] : 0.0, q2_km1 = moist ? pq2[qc] : 0.0;
            double om_km1 = omega_old(kc), om_k = omega_old(kc + 1);
            if (ka == 0) {        // face 0 and cell 0
                X[lane] = dtau * srz[q0];
                Z[lane] = fB[q0];
                omr[0] = om_km1;
                t0r[0] = t0_km1; t1r[0] = t1_km1;
            }
#pragma unroll
            for (int m = 0; m < CM; ++m) {
                const int k = ka + m;
                if (k == 0) continue;
                if (k > kb) break;
                const long q = q0 + (long)k * sz;
                const double Jk = J[q], s0_k = sr0[q], s1_k = sr1[q], p_k = prim[q], p_kp1 = prim[q + sz];
                const double drho_k = cc0[q] - sc0[q], drt_k = cc1[q] - sc1[q];
                const double q1_k = moist ? pq1[q] : 0.0, q2_k = moist ? pq2[q] : 0.0;
                const double fP_k = fP[q], fQ_k = fQ[q], srz_k = srz[q], fA_k = fA[q], fB_k = fB[q];
                const double dw_k = pw(k) - szm[q];
                const double om_kp1 = omega_old(k + 1);
                double t0_k, t1_k;
                trhs_at(k, t0_k, t1_k);
                // the slope terms of face k with the new increments: Omega of
                // them here, rho*w from the solution in phase 3 (kept, so the
                // increments and node heights are read once)
                double px_k, py_k;
                met::zface_slope_terms(i, j, k, fnu, fnv, z, dxi, dyi, px_k, py_k);
                const double ofw = (0. - px_k) - py_k;
                pxr[m] = px_k; pyr[m] = py_k;
                double detJ_on_kface = 0.5 * (Jk + J_km1);
                double coeff_P = fP_k;
                double coeff_Q = fQ_k;
                if (moist) {
                    double qf = 0.5 * (q1_k + q1_km1 + q2_k + q2_km1);
                    coeff_P /= (1.0 + qf);
                    coeff_Q /= (1.0 + qf);
                }
                double theta_t_lo = 0.5 * (p_km2 + p_km1);
                double theta_t_mid = 0.5 * (p_km1 + p_k);
                double theta_t_hi = 0.5 * (p_k + p_kp1);
                double R0_tmp = (-halfg * drho_k + coeff_P * drt_k) * Jk
                              + (-halfg * drho_km1 + coeff_Q * drt_km1) * J_km1;
                double R1_tmp = -halfg * (s0_k * Jk + s0_km1 * J_km1)
                              + (coeff_P * s1_k * Jk + coeff_Q * s1_km1 * J_km1);
                R1_tmp += (halfg) * (beta_1 * dzi * (om_kp1 - om_km1) + t0_k + t0_km1);
                R1_tmp += -(coeff_P * (beta_1 * dzi * (om_kp1 * theta_t_hi - om_k * theta_t_mid) + t1_k) +
                            coeff_Q * (beta_1 * dzi * (om_k * theta_t_mid - om_km1 * theta_t_lo) + t1_km1));
                double rhs = detJ_on_kface * dw_k + dtau * (detJ_on_kface * srz_k + R0_tmp + dtau * beta_2 * R1_tmp);
                rhs += detJ_on_kface * ofw;
                X[k * 64 + lane] = rhs;
                Y[k * 64 + lane] = fA_k;
                Z[k * 64 + lane] = fB_k;
                omr[m] = om_k;
                t0r[m] = t0_k; t1r[m] = t1_k;
                tile_level_fence();
                om_km1 = om_k; om_k = om_kp1;
                t0_km1 = t0_k; t1_km1 = t1_k;
                drho_km1 = drho_k; drt_km1 = drt_k;
                J_km1 = Jk; s0_km1 = s0_k; s1_km1 = s1_k;
                p_km2 = p_km1; p_km1 = p_k;
                q1_km1 = q1_k; q2_km1 = q2_k;
            }
#pragma unroll
            for (int m = 0; m <= CM; ++m)      // Omega of the face above the chunk
                if (m == kb - ka + 1) omr[m] = om_k;
            if (kb == nz - 1) {   // the top face: the right-hand side of its elimination step
                const long qt = q0 + (long)nz * sz;
                X[nz * 64 + lane] = dtau * srz[qt];
                Y[nz * 64 + lane] = fA[qt];
                Z[nz * 64 + lane] = fB[qt];
            }
        }
        __syncthreads();
        // ---- phase 2: elimination (wave 0), c_k / b_k and Omega staged, substitution (wave 0) ----
        if (wave == 0 && valid) {
            tile_forward(X, Y, Z, nz, lane);
            const long qt = q0 + (long)nz * sz;
            cz[qt] = szm[qt];
        }
        __syncthreads();
        if (work) {
#pragma unroll
            for (int m = 0; m < CM; ++m) {
                const int k = ka + m;
                if (k > kb) break;
                Y[k * 64 + lane] = fC[q0 + (long)k * sz] * Z[k * 64 + lane];
                Z[k * 64 + lane] = omr[m];
            }
            if (kb == nz - 1) {
#pragma unroll
                for (int m = 0; m <= CM; ++m)
                    if (m == kb - ka + 1) Z[nz * 64 + lane] = omr[m];
            }
        }
        __syncthreads();
        if (wave == 0 && valid) tile_backward(X, Y, nz, lane);
        __syncthreads();
        // ---- phase 3: rho*w, the averaged Omega, rho and rho*theta of the chunk ----
        if (work) {
#pragma unroll
            for (int m = 0; m < CM; ++m) {
                const int k = ka + m;
                if (k > kb) break;
                const long q = q0 + (long)k * sz;
                const double sol = X[k * 64 + lane], sol_hi = X[(k + 1) * 64 + lane];
                const double om_lo = Z[k * 64 + lane], om_hi = Z[(k + 1) * 64 + lane];
                const double szm_k = szm[q], azm_k = azm[q];
                const double Jq = J[q], pq = prim[q], p_lo = prim[q - sz], p_hi = prim[q + sz];
                const double sr0_k = sr0[q], sr1_k = sr1[q], cc0_k = cc0[q], cc1_k = cc1[q];
                const double cz_k = (k >= 1) ? szm_k + ((sol + pxr[m]) + pyr[m]) : szm_k;
                double zflux_lo = beta_2 * sol + beta_1 * om_lo;
                double zflux_hi = beta_2 * sol_hi + beta_1 * om_hi;
                double fast_rhs_rho = -(t0r[m] + (zflux_hi - zflux_lo) * dzi) / Jq;
                double fast_rhs_rhotheta = -(t1r[m] + 0.5 * (zflux_hi * (pq + p_hi) -
                                                             zflux_lo * (pq + p_lo)) * dzi) / Jq;
                cz[q] = cz_k;
                azm[q] = azm_k + facinv * zflux_lo / (1.0 * 1.0);
                if (k == nz - 1) azm[q + sz] += facinv * zflux_hi / (1.0 * 1.0);
                cc0[q] = cc0_k + dtau * (sr0_k + fast_rhs_rho);
                cc1[q] = cc1_k + dtau * (sr1_k + fast_rhs_rhotheta);
                tile_level_fence();
            }
        }
    };
    // a recorded (batched) launch takes the largest chunk, so every member of a
    // batch records the same kernel
    const int cm = tile_cm8(nz) == 0 ? 0 : (g_rec ? 12 : tile_cm8(nz));
    if (cm != 0 && P.col_variant == COL_TILE) {
        auto launch = [&](auto cmc) {
            launch_tiles<decltype(cmc)::value, 8>(0, nx - 1, 0, ny - 1, s,
                [=] __device__(int i, int j, bool valid, int lane, int wave, double* lds) {
                    tile(i, j, valid, lane, wave, lds, cmc);
                });
        };
        switch (cm) {
            case 4: launch(std::integral_constant<int, 4>{}); break;
            case 8: launch(std::integral_constant<int, 8>{}); break;
            default: launch(std::integral_constant<int, 12>{}); break;
        }
    } else
    launch_cols(0, nx - 1, 0, ny - 1, s, [=] __device__(int i, int j) {
        const long q0 = L.ix(i, j, 0);
        // Omega of the old increments on face kf (0 at the bottom, the rho*w
        // increment at the top, OmegaFromW between)
        // prev rho*w on interior face kf (round trip with the two cells' density
        // in the substep chain); the top face keeps what the last substep wrote
        auto pw = [&](int kf) -> double {
            const long q = q0 + (long)kf * sz;
            if (!rt) return pzm[q];
            if (kf >= nz) return wtop[L.ix(i, j, -L.ng)];
            const double d = prho[q] + prho[q - sz];
            const double w = pzm[q] * 2.0 / d;
            return w * 0.5 * d;
        };
        auto omega_old = [&](int kf) -> double {
            const long q = q0 + (long)kf * sz;
            if (kf == 0) return 0.;
            const double dwk = pw(kf) - szm[q];
            if (kf == nz) return dwk;
            return met::omega_from_w(i, j, kf, dwk, fdu, fdv, z, dxi, dyi);
        };
        // horizontal divergence of the new increments in cell k
        auto trhs_at = [&](int k, double& t0, double& t1) {
            const long q = q0 + (long)k * sz;
            double h_zeta_cc_xface_hi = 0.5 * dzi * (z(i + 1, j, k + 1) + z(i + 1, j + 1, k + 1) - z(i + 1, j, k) - z(i + 1, j + 1, k));
            double h_zeta_cc_xface_lo = 0.5 * dzi * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k) - z(i, j + 1, k));
            double h_zeta_cc_yface_hi = 0.5 * dzi * (z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1) - z(i, j + 1, k) - z(i + 1, j + 1, k));
            double h_zeta_cc_yface_lo = 0.5 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k) - z(i + 1, j, k));
            double xflux_lo = nu[q] * h_zeta_cc_xface_lo / 1.0;
            double xflux_hi = nu[q + 1] * h_zeta_cc_xface_hi / 1.0;
            double yflux_lo = nv[q] * h_zeta_cc_yface_lo / 1.0;
            double yflux_hi = nv[q + sy] * h_zeta_cc_yface_hi / 1.0;
            double mfsq = 1.0 * 1.0;
            t0 = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq;
            t1 = ((xflux_hi * (prim[q] + prim[q + 1]) - xflux_lo * (prim[q] + prim[q - 1])) * dxi * mfsq +
                  (yflux_hi * (prim[q] + prim[q + sy]) - yflux_lo * (prim[q] + prim[q - sy])) * dyi * mfsq) * 0.5;
        };
        // face 0 and cell 0
        double s_prev = (dtau * srz[q0]) * fB[q0];
        soln[q0] = s_prev;
        double om_km1 = 0.0, om_k = omega_old(1);
        om[q0] = om_km1;
        double t0_km1, t1_km1;
        trhs_at(0, t0_km1, t1_km1);
        trhs0[q0] = t0_km1;
        trhs1[q0] = t1_km1;
        double drho_km1 = cc0[q0] - sc0[q0], drt_km1 = cc1[q0] - sc1[q0];
        double J_km1 = J[q0], s0_km1 = sr0[q0], s1_km1 = sr1[q0];
        double p_km2 = prim[q0 - sz], p_km1 = prim[q0];
        double q1_km1 = moist ? pq1[q0] : 0.0, q2_km1 = moist ? pq2[q0] : 0.0;
        // Each level reads everything it needs before its first store (the
        // stores may alias the inputs as far as the compiler knows, so a load
        // after a store waits for another round trip to memory).
        for (int k = 1; k <= nz - 1; ++k) {
            const long q = q0 + (long)k * sz;
            const double Jk = J[q], s0_k = sr0[q], s1_k = sr1[q], p_k = prim[q], p_kp1 = prim[q + sz];
            const double drho_k = cc0[q] - sc0[q], drt_k = cc1[q] - sc1[q];
            const double q1_k = moist ? pq1[q] : 0.0, q2_k = moist ? pq2[q] : 0.0;
            const double fP_k = fP[q], fQ_k = fQ[q], fA_k = fA[q], fB_k = fB[q], srz_k = srz[q];
            const double dw_k = pw(k) - szm[q];
            const double om_kp1 = omega_old(k + 1);
            double t0_k, t1_k;
            trhs_at(k, t0_k, t1_k);
            const double ofw = met::omega_from_w(i, j, k, 0., fnu, fnv, z, dxi, dyi);
            double detJ_on_kface = 0.5 * (Jk + J_km1);
            double coeff_P = fP_k;
            double coeff_Q = fQ_k;
            if (moist) {
                double qf = 0.5 * (q1_k + q1_km1 + q2_k + q2_km1);
                coeff_P /= (1.0 + qf);
                coeff_Q /= (1.0 + qf);
            }
            double theta_t_lo = 0.5 * (p_km2 + p_km1);
            double theta_t_mid = 0.5 * (p_km1 + p_k);
            double theta_t_hi = 0.5 * (p_k + p_kp1);
            double R0_tmp = (-halfg * drho_k + coeff_P * drt_k) * Jk
                          + (-halfg * drho_km1 + coeff_Q * drt_km1) * J_km1;
            double R1_tmp = -halfg * (s0_k * Jk + s0_km1 * J_km1)
                          + (coeff_P * s1_k * Jk + coeff_Q * s1_km1 * J_km1);
            R1_tmp += (halfg) * (beta_1 * dzi * (om_kp1 - om_km1) + t0_k + t0_km1);
            R1_tmp += -(coeff_P * (beta_1 * dzi * (om_kp1 * theta_t_hi - om_k * theta_t_mid) + t1_k) +
                        coeff_Q * (beta_1 * dzi * (om_k * theta_t_mid - om_km1 * theta_t_lo) + t1_km1));
            double rhs = detJ_on_kface * dw_k + dtau * (detJ_on_kface * srz_k + R0_tmp + dtau * beta_2 * R1_tmp);
            rhs += detJ_on_kface * ofw;
            s_prev = (rhs - fA_k * s_prev) * fB_k;
            om[q] = om_k;
            trhs0[q] = t0_k;
            trhs1[q] = t1_k;
            soln[q] = s_prev;
            om_km1 = om_k; om_k = om_kp1;
            t0_km1 = t0_k; t1_km1 = t1_k;
            drho_km1 = drho_k; drt_km1 = drt_k;
            J_km1 = Jk; s0_km1 = s0_k; s1_km1 = s1_k;
            p_km2 = p_km1; p_km1 = p_k;
            q1_km1 = q1_k; q2_km1 = q2_k;
        }
        // top face: the last elimination step (no substitution there)
        const long qt = q0 + (long)nz * sz;
        double sol_hi = (dtau * srz[qt] - fA[qt] * s_prev) * fB[qt];
        cz[qt] = szm[qt];
        // one downward sweep: substitution, rho*w of face k, then the update
        // of cell k from its two faces
        double om_hi = om_k, p_hi = prim[qt];
        for (int k = nz - 1; k >= 0; --k) {
            const long q = q0 + k * sz;
            // loads first, as in the forward sweep
            const double soln_k = soln[q], fCB = fC[q] * fB[q], szm_k = szm[q], om_lo = om[q], azm_k = azm[q];
            const double Jq = J[q], pq = prim[q], p_lo = prim[q - sz], tr0 = trhs0[q], tr1 = trhs1[q];
            const double sr0_k = sr0[q], sr1_k = sr1[q], cc0_k = cc0[q], cc1_k = cc1[q];
            const double sol = soln_k - fCB * sol_hi;
            const double cz_k = (k >= 1) ? szm_k + met::w_from_omega(i, j, k, sol, fnu, fnv, z, dxi, dyi) : szm_k;
            double zflux_lo = beta_2 * sol + beta_1 * om_lo;
            double zflux_hi = beta_2 * sol_hi + beta_1 * om_hi;
            double fast_rhs_rho = -(tr0 + (zflux_hi - zflux_lo) * dzi) / Jq;
            double fast_rhs_rhotheta = -(tr1 + 0.5 * (zflux_hi * (pq + p_hi) -
                                                      zflux_lo * (pq + p_lo)) * dzi) / Jq;
            cz[q] = cz_k;
            azm[q] = azm_k + facinv * zflux_lo / (1.0 * 1.0);
            if (k == nz - 1) azm[q + sz] += facinv * zflux_hi / (1.0 * 1.0);
            cc0[q] = cc0_k + dtau * (sr0_k + fast_rhs_rho);
            cc1[q] = cc1_k + dtau * (sr1_k + fast_rhs_rhotheta);
            sol_hi = sol;
            om_hi = om_lo;
            p_hi = pq;
        }
    });
    // algorithmic bytes per cell: the fields of the substep this kernel reads
    // (stage rho, rho*theta, rho*u, rho*v, rho*w; the previous rho*u, rho*v,
    // rho*w; the new rho*u, rho*v; the current rho, rho*theta; the slow RHS of
    // rho, rho*theta, rho*w; the averaged rho*w; z_nd, detJ; qv, qc when moist)
    // and writes (rho*w, averaged rho*w, rho, rho*theta), 8 B each
    kt_stop(P.timer, KT_FAST_COLUMN, s, (moist ? 24.0 : 22.0) * 8.0 * nx * ny * nz, "fast_rhs_T column");
}

}  // namespace erfx
