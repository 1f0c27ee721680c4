// erfx_internal.h -- launchers shared between the library's translation units.
#pragma once

#include "erfx_common.h"

namespace erfx {

// erfx_bc.hip
void fill_boundary_periodic(const Params& P, double* a, int ncomp, int ng, int nodal_dir, hipStream_t s);
void fill_cons(const Params& P, double* c, int icomp, int ncomp, int ng, hipStream_t s);
void fill_xvel(const Params& P, double* a, int ng, hipStream_t s);
void fill_yvel(const Params& P, double* a, int ng, hipStream_t s);
// u, v: the filled velocities, read for the wall value of w on a terrain grid
void fill_zvel(const Params& P, double* a, int ng, hipStream_t s, const double* u = nullptr,
               const double* v = nullptr);
void momentum_to_velocity(const Params& P, double* xv, double* yv, double* zv, const double* dens,
                          const double* xm, const double* ym, const double* zm, hipStream_t s);
void velocity_to_momentum(const Params& P, const double* xv, const double* yv, const double* zv, int gx,
                          int gz_w, const double* dens, double* xm, double* ym, double* zm, hipStream_t s);
void fill_intermediate_patch(const Params& P, const erfx_state& S, double* xv, double* yv, double* zv,
                             int ngc, int ngv, bool cons_only, int icomp, int ncomp, hipStream_t s);
void apply_bcs(const Params& P, const erfx_state& S, double* xv, double* yv, double* zv, int ngc, int ngv,
               bool fast_only, bool synced, hipStream_t s);
void fill_patch(const Params& P, double* c, double* u, double* v, double* w, hipStream_t s);
void cons_to_prim(const Params& P, const double* c, double* prim, double* pi_stage, int ng, hipStream_t s);

// Device-to-device copy of up to COPY_MAX arrays in one launch over the whole
// chip (the runtime's copy engine splits a large copy into many small blit
// dispatches that reach a fraction of the HBM bandwidth).  n: doubles, even.
constexpr int COPY_MAX = 4;
struct CopyList {
    double* dst[COPY_MAX];
    const double* src[COPY_MAX];
    long n[COPY_MAX];
    int count = 0;
    void add(double* d, const double* s_, long n_) { dst[count] = d; src[count] = s_; n[count] = n_; ++count; }
};
void copy_fields(const CopyList& c, hipStream_t s);

// erfx_api.hip: the checks of erfx_ctx_create on a configuration (status code,
// erfx_last_error text); a context's run constants and stream
int validate_config(const erfx_config& c);
const Params& ctx_params(const erfx_ctx* c);
hipStream_t ctx_stream(const erfx_ctx* c);
// erfx_views.hip: free the staging fields of a context's view hooks
void views_release(const erfx_ctx* c);

// erfx_fast.hip
struct FastWork {
    double* extrap;   // theta extrapolated forward, cells grown by 1
    double* drw;      // delta rho*w at the previous substep, z-faces grown by 1 laterally
    double* trhs;     // 2 comps: flux divergence of rho and rho*theta increments, cells
    double* soln;     // tridiagonal solution on z-faces
};
void make_fast_coeffs(const Params& P, double* fc, const erfx_state& stage, const double* prim,
                      const double* pi_stage, const double* r0, const double* pi0, double dtau, double beta_s,
                      hipStream_t s);
// Options of the sim driver's substep chain; all off for the C-ABI hook.
struct FastChain {
    // prev is S_sum of the previous substep without its apply_bcs: apply the
    // momentum -> velocity -> momentum round trip of that apply_bcs on read
    bool rt_prev = false;
    // write ghost layer 1 of rho and rho*theta (periodic images, foextrap rows)
    bool ghost_writes = false;
    double w_lo = 0.0, w_hi = 0.0;   // Dirichlet w on the bottom and top faces
};
void fast_rhs_N(const Params& P, const FastWork& W, int step, int nrk, const erfx_state& slow_rhs,
                const erfx_state& prev, const erfx_state& stage, const double* prim, const double* pi_stage,
                const double* fc, const erfx_state& data, const erfx_state& scratch, double dtau,
                double beta_s, double facinv, hipStream_t s, const FastChain& ch = FastChain());
bool fast_bcs_fusable(const Params& P);
// no_substep_fun (substepping_type None): S_sum = S_old + slow_dt * F_slow for
// rho, rho*theta and the momenta
void no_substep(const Params& P, const erfx_state& ssum, const erfx_state& sold, const erfx_state& fslow,
                double slow_dt, hipStream_t s);
// erf_fast_rhs_T on a static terrain grid (erfx_terrain.hip): increments of the
// previous substep and the updated horizontal momenta on faces grown by one
struct TerrainWork {
    double *du, *dv;          // horizontal momentum increments prev - stage
    double *nu, *nv;          // new horizontal momentum increments
    double *om;               // Omega of the old increments on z-faces
    double *wtop;             // one plane (i, j) at L.ix(i, j, -ng): rho*w of the top face, substep chain
};
void fast_rhs_T(const Params& P, const FastWork& W, const TerrainWork& T, int step, int nrk,
                const erfx_state& slow_rhs, const erfx_state& prev, const erfx_state& stage, const double* prim,
                const double* pi_stage, const double* fc, const erfx_state& data, const erfx_state& scratch,
                double dtau, double beta_s, double facinv, hipStream_t s, const FastChain& ch = FastChain());
// make_J, make_areas, make_zcc (ERF_TerrainMetrics.cpp:583-715) on the whole
// layout: detJ, ax, ay, az = 1 and z_cc = 0 outside the boxes they are made on
// (and mrx, mry: the ratios h_xi / h_zeta, h_eta / h_zeta of the pressure
// gradient's metric term on the x- and y-faces of the valid box, 0 elsewhere)
void terrain_metrics(const Params& P, const double* z_nd, double* detJ, double* ax, double* ay, double* az,
                     double* z_cc, double* mrx, double* mry, hipStream_t s);

// time-step estimate: max over valid cells of the acoustic and the advective
// inverse time scales into host[0], host[1] (stream-ordered; sync before use)
constexpr int EST_BLOCKS = 1024;
void est_dt_maxima(const Params& P, const double* cons, const double* u, const double* v, const double* w,
                   double* partial, double* host, hipStream_t s);
// double-double volume sums of rho, rho*theta, RhoScalar: host[0..5] = (hi, lo) x 3
// (partial needs 6 * EST_BLOCKS + 6 doubles)
void volume_sums(const Params& P, const double* cons, double* partial, double* host, hipStream_t s);
// PlaneAverage rows of the moist buoyancy (ERF_make_buoyancy.cpp:216-255) into
// out[r * nz + k]: r = 0 rho, 1 theta, 2 qv, 3 qc, 4 qp (exactly rounded sums of
// value / (nx ny) over level k), and PLANE_T = the temperature of those averages.
constexpr int PLANE_ROWS = 5, PLANE_T = 5, PLANE_NOUT = 6;
constexpr int PLANE_SPLIT_MAX = 64;
// partial needs nz * PLANE_SPLIT_MAX * 2 * PLANE_ROWS doubles; out PLANE_NOUT * nz
void plane_averages(const Params& P, const double* cons, const double* prim, double* partial, double* out,
                    hipStream_t s);
// Kessler microphysics on cons (erfx_kessler.hip); fz: z-face scratch field,
// rain_accum: k = 0 plane or null; *bad (device) is set to 1 when a
// temperature falls below the saturation fit's range (ERF aborts there)
void kessler_advance(const Params& P, double* cons, double* fz, double* rain_accum, double dt, hipStream_t s,
                     int* bad);

// erfx_slow.hip
struct SlowWork {
    double* pp;                   // perturbational pressure, cells grown by 1
    double* tau[6];               // 11, 22, 33, 12, 13, 23
    double* xsrc; double* ysrc; double* zsrc;
    double* pavg = nullptr;       // moist buoyancy plane averages (PLANE_NOUT x nz)
    double* pavg_part = nullptr;  // their partial sums
    double* omega = nullptr;      // Omega on z-faces grown by one (terrain grids)
    // terrain grids with diffusion: the non-symmetric stresses tau21, tau31,
    // tau32, the expansion rate, the two lateral diffusive fluxes of one
    // scalar and the diffusion term of each diffused cons component
    double* tau_t[3] = {nullptr, nullptr, nullptr};
    double* er = nullptr;
    double* dfx = nullptr;
    double* dfy = nullptr;
    double* dstate = nullptr;     // ncons components (cons indexing)
};
// erf_make_tau_terms and DiffusionSrcForState_T on a terrain grid (molecular
// diffusion; erfx_slow.hip): stresses into W.tau / W.tau_t, and the diffusion
// term of cons components n0..n0+nc-1 into W.dstate
void make_tau_terms_T(const Params& P, const SlowWork& W, const double* cons, const double* u, const double* v,
                      const double* w, hipStream_t s);
void diffusion_state_T(const Params& P, const SlowWork& W, int n0, int nc, const double* cons, const double* prim,
                       hipStream_t s);
void turbulent_viscosity(const Params& P, const SlowWork& W, const double* cons, const double* u,
                         const double* v, const double* w, double* eddy, hipStream_t s);
void slow_rhs_pre(const Params& P, const SlowWork& W, int nrk, double slow_dt, const erfx_state& rhs,
                  const erfx_state& data, const double* prim, const erfx_state& scratch, const double* u,
                  const double* v, const double* w, const double* eddy, const double* r0, const double* p0,
                  hipStream_t s);
void slow_rhs_post(const Params& P, const SlowWork& W, int nrk, double slow_dt, const erfx_state& rhs,
                   const erfx_state& old,
                   const erfx_state& snew, const erfx_state& data, const double* prim,
                   const erfx_state& scratch, const double* eddy, hipStream_t s);

}  // namespace erfx
