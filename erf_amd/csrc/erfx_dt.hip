// erfx_dt.hip -- global reductions of the step:
//  * the two maxima of ERF::estTimeStep (Source/TimeIntegration/ERF_ComputeTimestep.cpp:66-163),
//    a two-pass max-reduction (a max is exact and independent of the order of
//    reduction, so the result equals the serial loop's bit for bit);
//  * the volume-weighted sums of ERF::volWgtSumMF (Source/IO/ERF_WriteScalarProfiles.cpp:378-410),
//    kept in double-double all the way, so the result is the exactly rounded
//    sum of the per-cell products whatever the reduction order;
//  * the plane averages of the moist buoyancy (PlaneAverage, Source/Utils/ERF_PlaneAverage.H:184-239),
//    per level, summed the same way.
#include "erfx_common.h"
#include "erfx_internal.h"

namespace erfx {

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// block max of two values into out[0..1]; blockDim.x == 256
__device__ __forceinline__ void block_max2(double a, double b, double* out) {
    __shared__ double sa[4], sb[4];
    a = wave_max(a);
    b = wave_max(b);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sa[w] = a; sb[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = fmax(fmax(sa[0], sa[1]), fmax(sa[2], sa[3]));
        out[1] = fmax(fmax(sb[0], sb[1]), fmax(sb[2], sb[3]));
    }
}

__global__ void __launch_bounds__(256) est_dt_pass1(Lay L, const double* cons, const double* u, const double* v,
                                                    const double* w, double dxi, double dyi, double dzi,
                                                    bool implicit, double* partial) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long n = (long)nx * ny * nz;
    double comp = -1.e100, lowM = -1.e100;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx);
        const int j = (int)((t / nx) % ny);
        const int k = (int)(t / ((long)nx * ny));
        const long q = L.ix(i, j, k);
        const double uc = 0.5 * (u[q] + u[q + 1]);
        const double vc = 0.5 * (v[q] + v[q + L.sy]);
        const double wc = 0.5 * (w[q] + w[q + L.sz]);
        const double rho = cons[q];
        const double rhotheta = cons[q + L.ne];
        const double pressure = eos_p_of_rt(rhotheta);
        const double c = sqrt(Gamma * pressure / rho);
        const double a = (fabs(uc) + c) * dxi;
        const double b = (fabs(vc) + c) * dyi;
        const double e = (fabs(wc) + c) * dzi;
        if (implicit) {
            if (nx > 1 && ny > 1) comp = fmax(fmax(a, b), comp);
            else if (nx > 1) comp = fmax(a, comp);
            else comp = fmax(b, comp);
        } else {
            if (nx > 1 && ny > 1) comp = fmax(fmax(fmax(a, b), e), comp);
            else if (nx > 1) comp = fmax(fmax(a, e), comp);
            else if (ny > 1) comp = fmax(fmax(b, e), comp);
            else comp = fmax(e, comp);
        }
        lowM = fmax(fmax(fmax(fabs(uc) * dxi, fabs(vc) * dyi), fabs(wc) * dzi), lowM);
    }
    block_max2(comp, lowM, partial + 2 * blockIdx.x);
}

__global__ void __launch_bounds__(256) est_dt_pass2(const double* partial, int nblocks, double* out) {
    double a = -1.e100, b = -1.e100;
    for (int t = threadIdx.x; t < nblocks; t += blockDim.x) {
        a = fmax(a, partial[2 * t]);
        b = fmax(b, partial[2 * t + 1]);
    }
    block_max2(a, b, out);
}

void est_dt_maxima(const Params& P, const double* cons, const double* u, const double* v, const double* w,
                   double* partial, double* host, hipStream_t s) {
    if (rec_refuse("time-step estimate")) return;
    hipLaunchKernelGGL(est_dt_pass1, dim3(EST_BLOCKS), dim3(256), 0, s, P.L, cons, u, v, w, P.dxInv[0], P.dxInv[1],
                       1.0 / P.dz_min, P.cfg.implicit_substepping != 0 && !P.cfg.no_substepping, partial);   // ERF_ComputeTimestep.cpp:66
    hipLaunchKernelGGL(est_dt_pass2, dim3(1), dim3(256), 0, s, (const double*)partial, EST_BLOCKS,
                       partial + 2 * EST_BLOCKS);
    (void)hipMemcpyAsync(host, partial + 2 * EST_BLOCKS, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
}

// ---- double-double sums ---------------------------------------------------------
struct DD { double hi, lo; };
__device__ __forceinline__ DD dd_add(DD a, DD b) {
    double s = a.hi + b.hi;
    double bb = s - a.hi;
    double e = (a.hi - (s - bb)) + (b.hi - bb);
    e += a.lo + b.lo;
    double h = s + e;
    return {h, e - (h - s)};
}
__device__ __forceinline__ DD dd_shfl(DD v, int off) { return {__shfl_xor(v.hi, off, 64), __shfl_xor(v.lo, off, 64)}; }

// block sum of NS double-double accumulators into out[2*c], out[2*c+1]
template <int NS>
__device__ __forceinline__ void block_sum(DD* acc, double* out) {
    __shared__ double sh[4][NS][2];
    for (int c = 0; c < NS; ++c)
        for (int off = 32; off > 0; off >>= 1) acc[c] = dd_add(acc[c], dd_shfl(acc[c], off));
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0)
        for (int c = 0; c < NS; ++c) { sh[w][c][0] = acc[c].hi; sh[w][c][1] = acc[c].lo; }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < NS; ++c) {
            DD t = {sh[0][c][0], sh[0][c][1]};
            for (int ww = 1; ww < 4; ++ww) t = dd_add(t, DD{sh[ww][c][0], sh[ww][c][1]});
            out[2 * c] = t.hi;
            out[2 * c + 1] = t.lo;
        }
}

// sums of (cons comp / mf^2) * cell volume over the valid cells for rho,
// rho*theta and RhoScalar (the quantities of ERF::sum_integrated_quantities);
// on a terrain grid the volume is cell_vol * detJ, made first as the
// reference's volume MultiFab is (ERF_WriteScalarProfiles.cpp:405-408)
__global__ void __launch_bounds__(256) sums_pass1(Lay L, const double* cons, const double* J, double cell_vol,
                                                  double* partial) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long n = (long)nx * ny * nz;
    const int comps[3] = {ERFX_RHO, ERFX_RHOTHETA, ERFX_RHOSCALAR};
    DD acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx);
        const int j = (int)((t / nx) % ny);
        const int k = (int)(t / ((long)nx * ny));
        const long q = L.ix(i, j, k);
        const double vol = J ? cell_vol * J[q] : cell_vol;
        for (int c = 0; c < 3; ++c) {
            double phi = cons[q + comps[c] * L.ne] / (1.0 * 1.0);
            acc[c] = dd_add(acc[c], DD{phi * vol, 0.0});
        }
    }
    block_sum<3>(acc, partial + 6 * blockIdx.x);
}

__global__ void __launch_bounds__(256) sums_pass2(const double* partial, int nblocks, double* out) {
    DD acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    for (int t = threadIdx.x; t < nblocks; t += blockDim.x)
        for (int c = 0; c < 3; ++c) acc[c] = dd_add(acc[c], DD{partial[6 * t + 2 * c], partial[6 * t + 2 * c + 1]});
    block_sum<3>(acc, out);
}

void volume_sums(const Params& P, const double* cons, double* partial, double* host, hipStream_t s) {
    if (rec_refuse("volume sums")) return;
    const double cell_vol = P.cfg.dx * P.cfg.dy * P.cfg.dz;
    hipLaunchKernelGGL(sums_pass1, dim3(EST_BLOCKS), dim3(256), 0, s, P.L, cons,
                       (const double*)(P.terrain ? P.detJ : nullptr), cell_vol, partial);
    hipLaunchKernelGGL(sums_pass2, dim3(1), dim3(256), 0, s, (const double*)partial, EST_BLOCKS,
                       partial + 6 * EST_BLOCKS);
    (void)hipMemcpyAsync(host, partial + 6 * EST_BLOCKS, 6 * sizeof(double), hipMemcpyDeviceToHost, s);
}

// ---- plane averages ------------------------------------------------------------
// work-group (b, k) sums the plane terms value * denom of its share of level k
__device__ __forceinline__ void plane_pass1(unsigned b, int k, Lay L, const double* cons, const double* prim,
                                            double denom, int nq, int nsplit, double* partial) {
    const int nx = L.nx, ny = L.ny;
    const long n = (long)nx * ny;
    DD acc[PLANE_ROWS];
    for (int r = 0; r < PLANE_ROWS; ++r) acc[r] = DD{0.0, 0.0};
    const double* pth = prim;
    const double* pq[3] = {prim + (ERFX_RHOQ1 - 1) * L.ne, prim + (ERFX_RHOQ2 - 1) * L.ne,
                           prim + (ERFX_RHOQ3 - 1) * L.ne};
    for (long t = (long)b * blockDim.x + threadIdx.x; t < n; t += (long)nsplit * blockDim.x) {
        const int i = (int)(t % nx), j = (int)(t / nx);
        const long q = L.ix(i, j, k);
        const double fac = 1.0;
        double v[PLANE_ROWS] = {cons[q], pth[q], 0.0, 0.0, 0.0};
        for (int c = 0; c < nq && c < 3; ++c) v[2 + c] = pq[c][q];
        for (int r = 0; r < PLANE_ROWS; ++r) acc[r] = dd_add(acc[r], DD{v[r] * denom * fac, 0.0});
    }
    block_sum<PLANE_ROWS>(acc, partial + ((long)k * nsplit + b) * 2 * PLANE_ROWS);
}

// block k folds the nsplit partials of level k, then the temperature of the
// averages (getTgivenRandRTh(rho_d, rho_d theta_d, qv_d), ERF_make_buoyancy.cpp:274-275)
__device__ __forceinline__ void plane_pass2(int k, const double* partial, int nsplit, int nz, double* out) {
    DD acc[PLANE_ROWS];
    for (int r = 0; r < PLANE_ROWS; ++r) acc[r] = DD{0.0, 0.0};
    const double* p = partial + (long)k * nsplit * 2 * PLANE_ROWS;
    for (int t = threadIdx.x; t < nsplit; t += blockDim.x)
        for (int r = 0; r < PLANE_ROWS; ++r)
            acc[r] = dd_add(acc[r], DD{p[t * 2 * PLANE_ROWS + 2 * r], p[t * 2 * PLANE_ROWS + 2 * r + 1]});
    __shared__ double res[2 * PLANE_ROWS];
    block_sum<PLANE_ROWS>(acc, res);
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int r = 0; r < PLANE_ROWS; ++r) out[r * nz + k] = res[2 * r];
        const double rho_d = res[0], theta_d = res[2], qv_d = res[4];
        out[PLANE_T * nz + k] = eos_T(rho_d, rho_d * theta_d, qv_d);
    }
}

void plane_averages(const Params& P, const double* cons, const double* prim, double* partial, double* out,
                    hipStream_t s) {
    const int nz = P.L.nz;
    const long nplane = (long)P.L.nx * P.L.ny;
    int nsplit = (2048 + nz - 1) / nz;
    if (nsplit > PLANE_SPLIT_MAX) nsplit = PLANE_SPLIT_MAX;
    const long need = (nplane + 255) / 256;
    if (nsplit > need) nsplit = (int)need;
    if (nsplit < 1) nsplit = 1;
    const double denom = 1.0 / (double)nplane;
    const Lay L = P.L;
    const int nq = P.n_qstate;
    launch_grid((unsigned)(nsplit * nz), s, [=] __device__(unsigned u) {
        plane_pass1(u % (unsigned)nsplit, (int)(u / (unsigned)nsplit), L, cons, prim, denom, nq, nsplit, partial);
    });
    const double* part = partial;
    launch_grid((unsigned)nz, s, [=] __device__(unsigned u) { plane_pass2((int)u, part, nsplit, nz, out); });
}

}  // namespace erfx
