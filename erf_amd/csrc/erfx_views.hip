// erfx_views.hip -- the hooks on Array4-shaped views: each field of a box as
// ERF allocates it (its own ghost width, index type and strides), for an
// adapter that forwards ERF's MultiFab data box by box.
//
// The library's kernels address every field of a box with one index (the
// common layout of include/erfx.h), which is what lets one fused kernel stream
// many fields.  A view in that layout is used in place.  Any other view is
// gathered into a staging field of the common layout before the hook and, if
// the hook writes it, scattered back after it: one launch per field each way,
// i-fastest on both sides.  Views of the same memory (ERF passes S_prev ==
// S_data to the substeps after the first) share one staging field, so the
// hook sees the aliasing the reference relies on.  The results are those of
// the pointer hooks, bit for bit.
#include "erfx_common.h"
#include "erfx_internal.h"

#include <map>
#include <string>
#include <vector>

namespace erfx {

struct ViewPool {
    std::vector<std::pair<size_t, double*>> free_;   // staging fields not in use: (doubles, pointer)
    std::vector<std::pair<size_t, double*>> all;
};
static std::map<const erfx_ctx*, ViewPool> g_pools;

void views_release(const erfx_ctx* c) {
    auto it = g_pools.find(c);
    if (it == g_pools.end()) return;
    for (auto& b : it->second.all) (void)hipFree(b.second);
    g_pools.erase(it);
}

// what a field of each role must be: components, index type, ghost width
struct RoleSpec {
    const char* name;
    int ncomp, nodal, ghost;   // nodal: -1 cell-centred, else the nodal direction
};
static int role_spec(const Params& P, int role, RoleSpec& r) {
    const int ngc = P.ng_cons, ngv = P.ng_vel, nc = P.cfg.ncons;
    switch (role) {
        case ERFX_ROLE_CONS: r = {"cons", nc, -1, ngc}; return 0;
        case ERFX_ROLE_XMOM: r = {"x-momentum", 1, 0, ngv}; return 0;
        case ERFX_ROLE_YMOM: r = {"y-momentum", 1, 1, ngv}; return 0;
        case ERFX_ROLE_ZMOM: r = {"z-momentum", 1, 2, ngv}; return 0;
        case ERFX_ROLE_XVEL: r = {"x-velocity", 1, 0, ngv}; return 0;
        case ERFX_ROLE_YVEL: r = {"y-velocity", 1, 1, ngv}; return 0;
        case ERFX_ROLE_ZVEL: r = {"z-velocity", 1, 2, ngv}; return 0;
        case ERFX_ROLE_PRIM: r = {"prim", nc - 1, -1, ngc}; return 0;
        case ERFX_ROLE_PI: r = {"pi_stage", 1, -1, 1}; return 0;
        case ERFX_ROLE_FAST_COEFFS: r = {"fast_coeffs", 5, 2, 0}; return 0;
        case ERFX_ROLE_EDDY: r = {"eddy diffusivities", 13, -1, 1}; return 0;
    }
    return -1;
}

static int check_view(const Params& P, int role, const erfx_view* v) {
    RoleSpec r;
    if (role_spec(P, role, r)) { set_error("unknown view role"); return ERFX_ERR_ARG; }
    if (!v || !v->p) { set_error(std::string("null view for ") + r.name); return ERFX_ERR_ARG; }
    const int n[3] = {P.L.nx, P.L.ny, P.L.nz};
    std::string why;
    if (v->ncomp < r.ncomp) why = "has " + std::to_string(v->ncomp) + " components, needs " + std::to_string(r.ncomp);
    for (int d = 0; d < 3 && why.empty(); ++d) {
        const int nd = (r.nodal == d) ? 1 : 0;
        if (v->ixtype[d] != nd) why = std::string("has the wrong index type along ") + "xyz"[d];
        else if (v->lo[d] > -r.ghost || v->hi[d] < n[d] - 1 + nd + r.ghost)
            why = std::string("does not cover the valid box grown by ") + std::to_string(r.ghost) + " along " + "xyz"[d];
    }
    if (why.empty()) {
        const long nxv = (long)v->hi[0] - v->lo[0] + 1, nyv = (long)v->hi[1] - v->lo[1] + 1;
        const long nzv = (long)v->hi[2] - v->lo[2] + 1;
        if (v->js < nxv || v->ks < v->js * nyv || (v->ncomp > 1 && v->ns < v->ks * nzv))
            why = "has strides smaller than its extents";
    }
    if (!why.empty()) {
        set_error(std::string("view for ") + r.name + " " + why);
        return ERFX_ERR_ARG;
    }
    return ERFX_OK;
}

// view element (i, j, k, n) <-> common layout, over the overlap of the view's
// box with the layout's (the valid box with ng ghosts and the extra face)
template <bool TO_LAYOUT>
static void copy_view(const Params& P, const erfx_view& v, int ncomp, double* lay, hipStream_t s) {
    const Lay L = P.L;
    const int g = L.ng;
    Box b(std::max(v.lo[0], -g), std::max(v.lo[1], -g), std::max(v.lo[2], -g), std::min(v.hi[0], L.nx + g),
          std::min(v.hi[1], L.ny + g), std::min(v.hi[2], L.nz + g));
    double* vp = v.p;
    const int l0 = v.lo[0], l1 = v.lo[1], l2 = v.lo[2];
    const long js = v.js, ks = v.ks, ns = v.ns;
    launch_box(b, s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        const long w = (long)(i - l0) + (long)(j - l1) * js + (long)(k - l2) * ks;
        for (int n = 0; n < ncomp; ++n) {
            if (TO_LAYOUT) lay[q + n * L.ne] = vp[w + n * ns];
            else vp[w + n * ns] = lay[q + n * L.ne];
        }
    });
}

// The common-layout pointers of one hook call: views staged in, and the
// outputs scattered back by finish().
class Staging {
  public:
    Staging(erfx_ctx* c, const Params& P, hipStream_t s) : c_(c), P_(P), s_(s) {}
    ~Staging() {
        ViewPool& pool = g_pools[c_];
        for (auto& e : ents_)
            if (e.buf && e.owned) pool.free_.push_back({e.size, e.buf});
    }
    // common-layout pointer standing for view v in role `role`
    int add(int role, const erfx_view* v, bool out, double** ptr) {
        int rc = check_view(P_, role, v);
        if (rc) return rc;
        RoleSpec r;
        role_spec(P_, role, r);
        for (auto& e : ents_)
            if (e.v.p == v->p) {
                if (e.v.js != v->js || e.v.ks != v->ks || e.v.ns != v->ns || e.v.lo[0] != v->lo[0] ||
                    e.v.lo[1] != v->lo[1] || e.v.lo[2] != v->lo[2]) {
                    set_error(std::string("two views of the same memory differ (") + r.name + ")");
                    return ERFX_ERR_ARG;
                }
                e.out = e.out || out;
                e.ncomp = std::max(e.ncomp, r.ncomp);
                *ptr = e.buf;
                return ERFX_OK;
            }
        Ent e;
        e.v = *v;
        e.out = out;
        e.ncomp = r.ncomp;
        const int g = P_.L.ng;
        const bool in_place = v->js == P_.L.sy && v->ks == P_.L.sz && (r.ncomp == 1 || v->ns == P_.L.ne) &&
                              v->lo[0] == -g && v->lo[1] == -g && v->lo[2] == -g;
        if (in_place) {
            e.buf = v->p;
            e.owned = false;
        } else {
            e.size = (size_t)r.ncomp * (size_t)P_.L.ne;
            rc = take(e.size, &e.buf);
            if (rc) return rc;
            e.owned = true;
            copy_view<true>(P_, *v, r.ncomp, e.buf, s_);
        }
        ents_.push_back(e);
        *ptr = e.buf;
        return ERFX_OK;
    }
    // state views: cons, xmom, ymom, zmom
    int add_state(const erfx_state_view* sv, bool out, erfx_state* st) {
        if (!sv) { set_error("null state view"); return ERFX_ERR_ARG; }
        int rc = add(ERFX_ROLE_CONS, &sv->cons, out, &st->cons);
        if (!rc) rc = add(ERFX_ROLE_XMOM, &sv->xmom, out, &st->xmom);
        if (!rc) rc = add(ERFX_ROLE_YMOM, &sv->ymom, out, &st->ymom);
        if (!rc) rc = add(ERFX_ROLE_ZMOM, &sv->zmom, out, &st->zmom);
        return rc;
    }
    // after the hook: scatter what it may have written back into the views
    int finish(int hook_rc) {
        if (hook_rc) return hook_rc;
        for (auto& e : ents_)
            if (e.out && e.owned) copy_view<false>(P_, e.v, e.ncomp, e.buf, s_);
        hipError_t err = hipGetLastError();
        if (err != hipSuccess) return hip_fail(err, "view scatter");
        return ERFX_OK;
    }

  private:
    struct Ent {
        erfx_view v;
        double* buf = nullptr;
        size_t size = 0;
        bool out = false, owned = false;
        int ncomp = 1;
    };
    int take(size_t n, double** p) {
        ViewPool& pool = g_pools[c_];
        for (size_t i = 0; i < pool.free_.size(); ++i)
            if (pool.free_[i].first == n) {
                *p = pool.free_[i].second;
                pool.free_.erase(pool.free_.begin() + (long)i);
                return ERFX_OK;
            }
        hipError_t e = hipMalloc((void**)p, n * sizeof(double));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc (view staging)");
        pool.all.push_back({n, *p});
        return ERFX_OK;
    }
    erfx_ctx* c_;
    const Params& P_;
    hipStream_t s_;
    std::vector<Ent> ents_;
};

}  // namespace erfx

using namespace erfx;

extern "C" {

int erfx_check_view(const erfx_config* cfg, int role, const erfx_view* v) {
    if (!cfg) { set_error("null argument"); return ERFX_ERR_ARG; }
    erfx_layout lay;
    int rc = erfx_layout_for(cfg, &lay);
    if (rc) return rc;
    Params P;
    make_params(*cfg, lay, P);
    return check_view(P, role, v);
}

int erfx_cons_to_prim_v(erfx_ctx* c, const erfx_view* cons, erfx_view* prim, erfx_view* pi_stage, int ng) {
    if (!c) { set_error("null argument"); return ERFX_ERR_ARG; }
    Staging S(c, ctx_params(c), ctx_stream(c));
    double *pc, *pp, *pi;
    int rc = S.add(ERFX_ROLE_CONS, cons, false, &pc);
    if (!rc) rc = S.add(ERFX_ROLE_PRIM, prim, true, &pp);
    if (!rc) rc = S.add(ERFX_ROLE_PI, pi_stage, true, &pi);
    if (rc) return rc;
    return S.finish(erfx_cons_to_prim(c, pc, pp, pi, ng));
}

int erfx_apply_bcs_v(erfx_ctx* c, erfx_state_view* state, erfx_view* xvel, erfx_view* yvel, erfx_view* zvel,
                     int ng_cons, int ng_vel, int fast_only, int vel_and_mom_synced) {
    if (!c) { set_error("null argument"); return ERFX_ERR_ARG; }
    Staging S(c, ctx_params(c), ctx_stream(c));
    erfx_state st;
    double *u, *v, *w;
    int rc = S.add_state(state, true, &st);
    if (!rc) rc = S.add(ERFX_ROLE_XVEL, xvel, true, &u);
    if (!rc) rc = S.add(ERFX_ROLE_YVEL, yvel, true, &v);
    if (!rc) rc = S.add(ERFX_ROLE_ZVEL, zvel, true, &w);
    if (rc) return rc;
    return S.finish(erfx_apply_bcs(c, &st, u, v, w, ng_cons, ng_vel, fast_only, vel_and_mom_synced));
}

int erfx_make_fast_coeffs_v(erfx_ctx* c, erfx_view* fast_coeffs, const erfx_state_view* stage,
                            const erfx_view* prim, const erfx_view* pi_stage, double dtau, double beta_s) {
    if (!c) { set_error("null argument"); return ERFX_ERR_ARG; }
    Staging S(c, ctx_params(c), ctx_stream(c));
    erfx_state st;
    double *fc, *pp, *pi;
    int rc = S.add(ERFX_ROLE_FAST_COEFFS, fast_coeffs, true, &fc);
    if (!rc) rc = S.add_state(stage, false, &st);
    if (!rc) rc = S.add(ERFX_ROLE_PRIM, prim, false, &pp);
    if (!rc) rc = S.add(ERFX_ROLE_PI, pi_stage, false, &pi);
    if (rc) return rc;
    return S.finish(erfx_make_fast_coeffs(c, fc, &st, pp, pi, dtau, beta_s));
}

static int fast_rhs_v(bool terrain, erfx_ctx* c, int step, int nrk, const erfx_state_view* slow_rhs,
                      const erfx_state_view* prev, const erfx_state_view* stage, const erfx_view* prim,
                      const erfx_view* pi_stage, const erfx_view* fast_coeffs, erfx_state_view* data,
                      erfx_state_view* scratch, double dtau, double beta_s, double facinv) {
    if (!c) { set_error("null argument"); return ERFX_ERR_ARG; }
    Staging S(c, ctx_params(c), ctx_stream(c));
    erfx_state sr, pv, sg, dt, sc;
    double *pp, *pi, *fc;
    // data and scratch first: prev aliases data after the first substep
    int rc = S.add_state(data, true, &dt);
    if (!rc) rc = S.add_state(scratch, true, &sc);
    if (!rc) rc = S.add_state(slow_rhs, false, &sr);
    if (!rc) rc = S.add_state(prev, false, &pv);
    if (!rc) rc = S.add_state(stage, false, &sg);
    if (!rc) rc = S.add(ERFX_ROLE_PRIM, prim, false, &pp);
    if (!rc) rc = S.add(ERFX_ROLE_PI, pi_stage, false, &pi);
    if (!rc) rc = S.add(ERFX_ROLE_FAST_COEFFS, fast_coeffs, false, &fc);
    if (rc) return rc;
    if (terrain)
        return S.finish(erfx_fast_rhs_T(c, step, nrk, &sr, &pv, &sg, pp, pi, fc, &dt, &sc, dtau, beta_s, facinv));
    return S.finish(erfx_fast_rhs_N(c, step, nrk, &sr, &pv, &sg, pp, pi, fc, &dt, &sc, dtau, beta_s, facinv));
}

int erfx_fast_rhs_N_v(erfx_ctx* c, int step, int nrk, const erfx_state_view* slow_rhs, const erfx_state_view* prev,
                      const erfx_state_view* stage, const erfx_view* prim, const erfx_view* pi_stage,
                      const erfx_view* fast_coeffs, erfx_state_view* data, erfx_state_view* scratch, double dtau,
                      double beta_s, double facinv) {
    return fast_rhs_v(false, c, step, nrk, slow_rhs, prev, stage, prim, pi_stage, fast_coeffs, data, scratch, dtau,
                      beta_s, facinv);
}

int erfx_fast_rhs_T_v(erfx_ctx* c, int step, int nrk, const erfx_state_view* slow_rhs, const erfx_state_view* prev,
                      const erfx_state_view* stage, const erfx_view* prim, const erfx_view* pi_stage,
                      const erfx_view* fast_coeffs, erfx_state_view* data, erfx_state_view* scratch, double dtau,
                      double beta_s, double facinv) {
    return fast_rhs_v(true, c, step, nrk, slow_rhs, prev, stage, prim, pi_stage, fast_coeffs, data, scratch, dtau,
                      beta_s, facinv);
}

int erfx_turbulent_viscosity_v(erfx_ctx* c, const erfx_view* cons, const erfx_view* xvel, const erfx_view* yvel,
                               const erfx_view* zvel, erfx_view* eddy) {
    if (!c) { set_error("null argument"); return ERFX_ERR_ARG; }
    Staging S(c, ctx_params(c), ctx_stream(c));
    double *pc, *u, *v, *w, *e;
    int rc = S.add(ERFX_ROLE_CONS, cons, false, &pc);
    if (!rc) rc = S.add(ERFX_ROLE_XVEL, xvel, false, &u);
    if (!rc) rc = S.add(ERFX_ROLE_YVEL, yvel, false, &v);
    if (!rc) rc = S.add(ERFX_ROLE_ZVEL, zvel, false, &w);
    if (!rc) rc = S.add(ERFX_ROLE_EDDY, eddy, true, &e);
    if (rc) return rc;
    return S.finish(erfx_turbulent_viscosity(c, pc, u, v, w, e, nullptr));
}

int erfx_slow_rhs_pre_v(erfx_ctx* c, int nrk, double slow_dt, erfx_state_view* rhs, const erfx_state_view* data,
                        const erfx_view* prim, erfx_state_view* scratch, const erfx_view* xvel,
                        const erfx_view* yvel, const erfx_view* zvel, const erfx_view* eddy) {
    if (!c) { set_error("null argument"); return ERFX_ERR_ARG; }
    Staging S(c, ctx_params(c), ctx_stream(c));
    erfx_state sr, sd, sc;
    double *pp, *u, *v, *w, *e = nullptr;
    int rc = S.add_state(rhs, true, &sr);
    if (!rc) rc = S.add_state(scratch, true, &sc);
    if (!rc) rc = S.add_state(data, false, &sd);
    if (!rc) rc = S.add(ERFX_ROLE_PRIM, prim, false, &pp);
    if (!rc) rc = S.add(ERFX_ROLE_XVEL, xvel, false, &u);
    if (!rc) rc = S.add(ERFX_ROLE_YVEL, yvel, false, &v);
    if (!rc) rc = S.add(ERFX_ROLE_ZVEL, zvel, false, &w);
    if (!rc && eddy) rc = S.add(ERFX_ROLE_EDDY, eddy, false, &e);
    if (rc) return rc;
    return S.finish(erfx_slow_rhs_pre(c, nrk, slow_dt, &sr, &sd, pp, &sc, u, v, w, e, nullptr));
}

int erfx_slow_rhs_post_v(erfx_ctx* c, int nrk, double slow_dt, erfx_state_view* rhs, const erfx_state_view* old,
                         erfx_state_view* snew, erfx_state_view* data, const erfx_view* prim,
                         const erfx_state_view* scratch, const erfx_view* eddy) {
    if (!c) { set_error("null argument"); return ERFX_ERR_ARG; }
    Staging S(c, ctx_params(c), ctx_stream(c));
    erfx_state sr, so, sn, sd, sc;
    double *pp, *e = nullptr;
    int rc = S.add_state(rhs, true, &sr);
    if (!rc) rc = S.add_state(snew, true, &sn);
    if (!rc) rc = S.add_state(data, true, &sd);
    if (!rc) rc = S.add_state(old, false, &so);
    if (!rc) rc = S.add_state(scratch, false, &sc);
    if (!rc) rc = S.add(ERFX_ROLE_PRIM, prim, false, &pp);
    if (!rc && eddy) rc = S.add(ERFX_ROLE_EDDY, eddy, false, &e);
    if (rc) return rc;
    return S.finish(erfx_slow_rhs_post(c, nrk, slow_dt, &sr, &so, &sn, &sd, pp, &sc, nullptr, nullptr, e));
}

int erfx_kessler_advance_v(erfx_ctx* c, erfx_view* cons, double dt) {
    if (!c) { set_error("null argument"); return ERFX_ERR_ARG; }
    Staging S(c, ctx_params(c), ctx_stream(c));
    double* pc;
    int rc = S.add(ERFX_ROLE_CONS, cons, true, &pc);
    if (rc) return rc;
    rc = erfx_kessler_advance(c, pc, dt, nullptr);
    return S.finish(rc);
}

}  // extern "C"
