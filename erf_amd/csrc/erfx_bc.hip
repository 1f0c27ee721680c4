// erfx_bc.hip -- ghost-cell fills, momentum <-> velocity conversion and
// cons_to_prim on the GPU (level 0, one box).
//
// The fills keep ERF's order -- periodic copy, lateral x, lateral y, vertical --
// because corner ghosts read values written by the earlier stages
// (ERF_PhysBCFunct.cpp:16-367).  Each stage is one launch over its ghost slabs.
#include "erfx_common.h"
#include "erfx_internal.h"
#include "erfx_terrain.h"

namespace erfx {

// FabArray::FillBoundary for one periodic box: ghost points whose periodic
// image lies in the valid region take that value.  Launched over the lateral
// ghost slabs of the periodic directions only (z is never periodic here).
void fill_boundary_periodic(const Params& P, double* a, int ncomp, int ng, int nodal_dir, hipStream_t s) {
    const Lay L = P.L;
    const int n0 = L.nx, n1 = L.ny, n2 = L.nz;
    const int vh0 = n0 - 1 + (nodal_dir == 0), vh1 = n1 - 1 + (nodal_dir == 1), vh2 = n2 - 1 + (nodal_dir == 2);
    const bool px = P.periodic[0], py = P.periodic[1];
    if (!px && !py) return;
    const long ne = L.ne;
    // ghost layers of z are not periodic: restrict to valid k
    if (px) {
        // i' in [0, 2ng): lo slab then hi slab
        Box b(0, -ng, 0, 2 * ng - 1, vh1 + ng, vh2);
        launch_box(b, s, [=] __device__(int ip, int j, int k) {
            int i = ip < ng ? ip - ng : vh0 + 1 + (ip - ng);
            int qi = i < 0 ? i + n0 : i - n0;
            int qj = j;
            if (py) { if (qj < 0) qj += n1; else if (qj > vh1) qj -= n1; }
            if (qi < 0 || qi > vh0 || qj < 0 || qj > vh1) return;
            long d = L.ix(i, j, k), sidx = L.ix(qi, qj, k);
            for (int c = 0; c < ncomp; ++c) a[d + c * ne] = a[sidx + c * ne];
        });
    }
    if (py) {
        Box b(-ng, 0, 0, vh0 + ng, 2 * ng - 1, vh2);
        launch_box(b, s, [=] __device__(int i, int jp, int k) {
            int j = jp < ng ? jp - ng : vh1 + 1 + (jp - ng);
            int qj = j < 0 ? j + n1 : j - n1;
            int qi = i;
            if (px) { if (qi < 0) qi += n0; else if (qi > vh0) qi -= n0; }
            if (qi < 0 || qi > vh0 || qj < 0 || qj > vh1) return;
            long d = L.ix(i, j, k), sidx = L.ix(qi, qj, k);
            for (int c = 0; c < ncomp; ++c) a[d + c * ne] = a[sidx + c * ne];
        });
    }
}

// ERFPhysBCFunct_cons::operator() with do_fb: FillBoundary of all components,
// then impose_lateral_cons_bcs (ERF_BoundaryConditions_cons.cpp:17-246) and
// impose_vertical_cons_bcs (:261-494) on components [icomp, icomp+ncomp).
// the ext_dir values of every cons component on face `ori`
struct ExtVals {
    double v[ERFX_NCONS_MOIST];
};
static ExtVals ext_vals(const Params& P, int ori) {
    ExtVals e;
    for (int n = 0; n < ERFX_NCONS_MOIST; ++n) e.v[n] = P.cons_extdir[n][ori];
    return e;
}

void fill_cons(const Params& P, double* c, int icomp, int ncomp, int ng, hipStream_t s) {
    fill_boundary_periodic(P, c, P.cfg.ncons, P.ng_cons, -1, s);
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne;
    if (!P.periodic[0]) {
        const int tl = P.bck[3][0], th = P.bck[3][3];
        Box b(0, -ng, 0, 2 * ng - 1, ny - 1 + ng, nz - 1);
        ExtVals xl = ext_vals(P, 0), xh = ext_vals(P, 3);
        launch_box(b, s, [=] __device__(int ip, int j, int k) {
            bool lo = ip < ng;
            int i = lo ? ip - ng : nx + (ip - ng);
            int t = lo ? tl : th;
            int src;
            if (t == EXT_DIR) {   // ERF_BoundaryConditions_cons.cpp:68-104
                const long d = L.ix(i, j, k);
                for (int n = icomp; n < icomp + ncomp; ++n) c[d + n * ne] = lo ? xl.v[n] : xh.v[n];
                return;
            }
            if (t == FOEXTRAP) src = lo ? 0 : nx - 1;
            else if (t == REFLECT_EVEN) src = lo ? -1 - i : 2 * (nx - 1) + 1 - i;
            else return;
            long d = L.ix(i, j, k), sidx = L.ix(src, j, k);
            for (int n = icomp; n < icomp + ncomp; ++n) c[d + n * ne] = c[sidx + n * ne];
        });
    }
    if (!P.periodic[1]) {
        const int tl = P.bck[3][1], th = P.bck[3][4];
        Box b(-ng, 0, 0, nx - 1 + ng, 2 * ng - 1, nz - 1);
        ExtVals yl = ext_vals(P, 1), yh = ext_vals(P, 4);
        launch_box(b, s, [=] __device__(int i, int jp, int k) {
            bool lo = jp < ng;
            int j = lo ? jp - ng : ny + (jp - ng);
            int t = lo ? tl : th;
            int src;
            if (t == EXT_DIR) {
                const long d = L.ix(i, j, k);
                for (int n = icomp; n < icomp + ncomp; ++n) c[d + n * ne] = lo ? yl.v[n] : yh.v[n];
                return;
            }
            if (t == FOEXTRAP) src = lo ? 0 : ny - 1;
            else if (t == REFLECT_EVEN) src = lo ? -1 - j : 2 * (ny - 1) + 1 - j;
            else return;
            long d = L.ix(i, j, k), sidx = L.ix(i, src, k);
            for (int n = icomp; n < icomp + ncomp; ++n) c[d + n * ne] = c[sidx + n * ne];
        });
    }
    {   // vertical: foextrap on both z faces for every supported face type
        Box b(-ng, -ng, 0, nx - 1 + ng, ny - 1 + ng, 2 * ng - 1);
        launch_box(b, s, [=] __device__(int i, int j, int kp) {
            bool lo = kp < ng;
            int k = lo ? kp - ng : nz + (kp - ng);
            int src = lo ? 0 : nz - 1;
            long d = L.ix(i, j, k), sidx = L.ix(i, j, src);
            for (int n = icomp; n < icomp + ncomp; ++n) c[d + n * ne] = c[sidx + n * ne];
        });
    }
}

// Tangential/normal face-velocity fills along one lateral or vertical
// direction.  `dir` is the fill direction, `nodal` whether the field is
// staggered along it.  Rules follow ERF_BoundaryConditions_{x,y,z}vel.cpp.
struct FaceFill {
    int tl, th;       // kinds on lo/hi sides
    double vl, vh;    // Dirichlet values
};

__device__ __forceinline__ double vel_rule_cell(const double* a, const Lay& L, int t, double v, bool lo,
                                                int dir, int i, int j, int k, int n) {
    // cell-centred along dir (tangential component): ghost m in [-ng,-1] or [n, n+ng-1]
    int p[3] = {i, j, k};
    int m = p[dir];
    int q = p[dir];
    double sgn = 1.0;
    if (t == EXT_DIR) return v;
    if (t == FOEXTRAP) q = lo ? 0 : n - 1;
    else if (t == REFLECT_EVEN) q = lo ? -1 - m : 2 * (n - 1) + 1 - m;
    else if (t == REFLECT_ODD) { q = lo ? -1 - m : 2 * (n - 1) + 1 - m; sgn = -1.0; }
    else return NAN;
    p[dir] = q;
    double val = a[L.ix(p[0], p[1], p[2])];
    return sgn < 0 ? -val : val;
}

// ERFPhysBCFunct_u/v::operator() for the velocity component `comp` (0 = x, 1 = y).
static void fill_hvel(const Params& P, double* a, int comp, int ng, hipStream_t s) {
    fill_boundary_periodic(P, a, 1, P.ng_vel, comp, s);
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int nn[3] = {nx, ny, nz};
    // extents of the staggered box
    const int hx = nx - 1 + (comp == 0), hy = ny - 1 + (comp == 1);
    for (int dir = 0; dir < 2; ++dir) {
        if (P.periodic[dir]) continue;
        const int tl = P.bck[comp][dir], th = P.bck[comp][dir + 3];
        const double vl = P.vel_extdir[comp][dir], vh = P.vel_extdir[comp][dir + 3];
        const int n = nn[dir];
        if (dir == comp) {
            // normal component: ghosts beyond the boundary faces plus the faces
            // themselves for ext_dir / neumann_int (ERF_BoundaryConditions_xvel.cpp:55-110)
            Box b = dir == 0 ? Box(0, -ng, 0, 2 * ng + 1, hy + ng, nz - 1) : Box(-ng, 0, 0, hx + ng, 2 * ng + 1, nz - 1);
            launch_box(b, s, [=] __device__(int i0, int j0, int k) {
                int pp = dir == 0 ? i0 : j0;
                bool lo = pp <= ng;              // ng ghosts + the lo face
                int m = lo ? pp - ng : n + (pp - ng - 1);   // lo: -ng..0 ; hi: n..n+ng
                int t = lo ? tl : th;
                double v = lo ? vl : vh;
                int p[3] = {i0, j0, k};
                p[dir] = m;
                bool face = lo ? (m == 0) : (m == n);
                auto at = [&](int q) { int r[3] = {p[0], p[1], p[2]}; r[dir] = q; return a[L.ix(r[0], r[1], r[2])]; };
                double val;
                if (face) {
                    if (t == EXT_DIR) val = v;
                    else if (t == NEUMANN_INT) val = lo ? (4.0 * at(1) - at(2)) / 3.0 : (4.0 * at(n - 1) - at(n - 2)) / 3.0;
                    else return;
                } else {
                    if (t == EXT_DIR) val = v;
                    else if (t == FOEXTRAP) val = at(lo ? 0 : n);
                    else if (t == REFLECT_EVEN) val = at(lo ? -m : 2 * n - m);
                    else if (t == REFLECT_ODD) val = -at(lo ? -m : 2 * n - m);
                    else if (t == NEUMANN_INT) val = lo ? (4.0 * at(1) - at(2)) / 3.0 : (4.0 * at(n - 1) - at(n - 2)) / 3.0;
                    else return;
                }
                a[L.ix(p[0], p[1], p[2])] = val;
            });
        } else {
            Box b = dir == 0 ? Box(0, -ng, 0, 2 * ng - 1, hy + ng, nz - 1) : Box(-ng, 0, 0, hx + ng, 2 * ng - 1, nz - 1);
            launch_box(b, s, [=] __device__(int i0, int j0, int k) {
                int pp = dir == 0 ? i0 : j0;
                bool lo = pp < ng;
                int m = lo ? pp - ng : n + (pp - ng);
                int p[3] = {i0, j0, k};
                p[dir] = m;
                double val = vel_rule_cell(a, L, lo ? tl : th, lo ? vl : vh, lo, dir, p[0], p[1], p[2], n);
                a[L.ix(p[0], p[1], p[2])] = val;
            });
        }
    }
    {   // vertical, over the full grown box (ERF_BoundaryConditions_xvel.cpp:200-245)
        const int tl = P.bck[comp][2], th = P.bck[comp][5];
        const double vl = P.vel_extdir[comp][2], vh = P.vel_extdir[comp][5];
        Box b(-ng, -ng, 0, hx + ng, hy + ng, 2 * ng - 1);
        launch_box(b, s, [=] __device__(int i, int j, int kp) {
            bool lo = kp < ng;
            int k = lo ? kp - ng : nz + (kp - ng);
            double val = vel_rule_cell(a, L, lo ? tl : th, lo ? vl : vh, lo, 2, i, j, k, nz);
            a[L.ix(i, j, k)] = val;
        });
    }
}

void fill_xvel(const Params& P, double* a, int ng, hipStream_t s) { fill_hvel(P, a, 0, ng, s); }
void fill_yvel(const Params& P, double* a, int ng, hipStream_t s) { fill_hvel(P, a, 1, ng, s); }

// ERFPhysBCFunct_w::operator() (ERF_PhysBCFunct.cpp:271-367) and
// ERF_BoundaryConditions_zvel.cpp:16-268: the box is clipped to k in [0, nz];
// w(k=0) always takes the zlo Dirichlet value, which on a terrain grid is
// Omega's, w following the slope of the filled u, v (:220-257).
void fill_zvel(const Params& P, double* a, int ng, hipStream_t s, const double* uf, const double* vf) {
    fill_boundary_periodic(P, a, 1, P.ng_vel, 2, s);
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int nn[2] = {nx, ny};
    for (int dir = 0; dir < 2; ++dir) {
        if (P.periodic[dir]) continue;
        const int tl = P.bck[2][dir], th = P.bck[2][dir + 3];
        const double vl = P.vel_extdir[2][dir], vh = P.vel_extdir[2][dir + 3];
        const int n = nn[dir];
        Box b = dir == 0 ? Box(0, -ng, 0, 2 * ng - 1, ny - 1 + ng, nz) : Box(-ng, 0, 0, nx - 1 + ng, 2 * ng - 1, nz);
        launch_box(b, s, [=] __device__(int i0, int j0, int k) {
            int pp = dir == 0 ? i0 : j0;
            bool lo = pp < ng;
            int m = lo ? pp - ng : n + (pp - ng);
            int p[3] = {i0, j0, k};
            p[dir] = m;
            double val = vel_rule_cell(a, L, lo ? tl : th, lo ? vl : vh, lo, dir, p[0], p[1], p[2], n);
            a[L.ix(p[0], p[1], p[2])] = val;
        });
    }
    const double vl = P.vel_extdir[2][2], vh = P.vel_extdir[2][5];
    const int th = P.bck[2][5];
    // both planes in one launch: kp = 0 -> k = 0, kp = 1 -> k = nz
    Box b(-ng, -ng, 0, nx - 1 + ng, ny - 1 + ng, 1);
    if (P.terrain) {
        const Zn z{P.z_nd, L};
        const Fld u{uf, L}, v{vf, L};
        const double dxi = P.dxInv[0], dyi = P.dxInv[1];
        launch_box(b, s, [=] __device__(int i, int j, int kp) {
            if (kp == 0) { a[L.ix(i, j, 0)] = met::w_from_omega(i, j, 0, vl, u, v, z, dxi, dyi); return; }
            if (th == EXT_DIR) a[L.ix(i, j, nz)] = met::w_from_omega(i, j, nz, vh, u, v, z, dxi, dyi);
            else if (th == NEUMANN_INT) a[L.ix(i, j, nz)] = (4.0 * a[L.ix(i, j, nz - 1)] - a[L.ix(i, j, nz - 2)]) / 3.0;
        });
        return;
    }
    launch_box(b, s, [=] __device__(int i, int j, int kp) {
        if (kp == 0) { a[L.ix(i, j, 0)] = vl; return; }
        if (th == EXT_DIR) a[L.ix(i, j, nz)] = vh;
        else if (th == NEUMANN_INT) a[L.ix(i, j, nz)] = (4.0 * a[L.ix(i, j, nz - 1)] - a[L.ix(i, j, nz - 2)]) / 3.0;
    });
}

// MomentumToVelocity (Source/Utils/ERF_MomentumToVelocity.cpp:24-96) on the
// valid faces; one launch covers the three staggered boxes.
void momentum_to_velocity(const Params& P, double* xv, double* yv, double* zv, const double* dens,
                          const double* xm, const double* ym, const double* zm, hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    Box b(0, 0, 0, nx, ny, nz);
    launch_box(b, s, [=] __device__(int i, int j, int k) {
        long c = L.ix(i, j, k);
        if (j < ny && k < nz) xv[c] = xm[c] * 2.0 / (dens[c] + dens[c - 1]);
        if (i < nx && k < nz) yv[c] = ym[c] * 2.0 / (dens[c] + dens[c - L.sy]);
        if (i < nx && j < ny) zv[c] = zm[c] * 2.0 / (dens[c] + dens[c - L.sz]);
    });
}

// VelocityToMomentum (Source/Utils/ERF_VelocityToMomentum.cpp:24-114): faces
// of the valid box grown by g (w limited to k in [0, nz]).
void velocity_to_momentum(const Params& P, const double* xv, const double* yv, const double* zv, int gx,
                          int gz_w, const double* dens, double* xm, double* ym, double* zm, hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int zlo_w = gz_w > 0 ? 0 : 0, zhi_w = nz;
    Box b(-gx, -gx, -gx, nx + gx, ny + gx, nz + gx);
    launch_box(b, s, [=] __device__(int i, int j, int k) {
        long c = L.ix(i, j, k);
        if (j <= ny - 1 + gx && k <= nz - 1 + gx) xm[c] = xv[c] * 0.5 * (dens[c] + dens[c - 1]);
        if (i <= nx - 1 + gx && k <= nz - 1 + gx) ym[c] = yv[c] * 0.5 * (dens[c] + dens[c - L.sy]);
        if (i <= nx - 1 + gx && j <= ny - 1 + gx && k >= zlo_w && k <= zhi_w && k >= -gz_w && k <= nz + gz_w)
            zm[c] = zv[c] * 0.5 * (dens[c] + dens[c - L.sz]);
    });
}

// apply_bcs with unsynced momenta on a box periodic in x and y with wall /
// outflow z faces, in two launches instead of the ~30 of the generic sequence
// below.  It writes exactly the values that sequence leaves behind
// (FillIntermediatePatch for rho first, then M2V on the valid faces, the cons
// fills of the remaining components, the velocity fills, V2M on the box grown
// by 1 and FillBoundary of the momenta), point for point:
//   KA  cons ghosts and every velocity the sequence defines, each computed
//       from valid data (periodic images, same-column values), so no lane
//       reads a ghost another lane writes;
//   KB  momenta: V2M where the sequence applies it, the periodic image's V2M
//       on the FillBoundary layers.
// Points the sequence does not touch are not written.
bool fast_bcs_fusable(const Params& P) {
    if (!P.periodic[0] || !P.periodic[1]) return false;
    const int G = P.ng_cons > P.ng_vel ? P.ng_cons : P.ng_vel;
    if (P.L.nx < G + 1 || P.L.ny < G + 1 || P.L.nz < 2 * G) return false;
    for (int v = 0; v < 2; ++v)
        for (int o : {2, 5})
            if (P.bck[v][o] != EXT_DIR && P.bck[v][o] != FOEXTRAP) return false;
    if (P.bck[2][5] != EXT_DIR && P.bck[2][5] != NEUMANN_INT) return false;
    return P.cfg.ncons <= 8;
}

struct ConsRows {
    int r[8];   // foextrap rows filled below and above, per component (lateral reach the same)
};

void apply_bcs_periodic(const Params& P, const erfx_state& S, double* xv, double* yv, double* zv, int ngc,
                        int ngv, bool fast_only, hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int ngC = P.ng_cons, ngV = P.ng_vel;
    const int G = ngC > ngV ? ngC : ngV;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const int nc = P.cfg.ncons;
    // rows per component: rho by the first FillIntermediatePatch, the others by the second
    ConsRows R;
    const int ngu1 = ngc > ngv + 1 ? ngc : ngv + 1;
    const int ncomp2 = fast_only ? 1 : nc - 1;
    int rmax = 0;
    for (int n = 0; n < 8; ++n) {
        R.r[n] = n == 0 ? ngu1 : (n < 1 + ncomp2 && n < nc ? ngc : 0);
        if (R.r[n] > rmax) rmax = R.r[n];
    }
    double* c = S.cons;
    double* xm = S.xmom;
    double* ym = S.ymom;
    double* zm = S.zmom;
    const int xtl = P.bck[0][2], xth = P.bck[0][5], ytl = P.bck[1][2], yth = P.bck[1][5];
    const double xvl = P.vel_extdir[0][2], xvh = P.vel_extdir[0][5];
    const double yvl = P.vel_extdir[1][2], yvh = P.vel_extdir[1][5];
    const double zvl = P.vel_extdir[2][2], zvh = P.vel_extdir[2][5];
    const int zth = P.bck[2][5];
    const int kz = rmax > ngv ? rmax : (ngv > 1 ? ngv : 1);
    const Box grownA(-G, -G, -kz, nx + G, ny + G, nz - 1 + kz);
    const bool terr = P.terrain;
    const Zn zt{P.z_nd, L};
    const double dxi = P.dxInv[0], dyi = P.dxInv[1];
    const Box grownB(-G, -G, -1, nx + G, ny + G, nz);

    launch_box(grownA, s, [=] __device__(int i, int j, int k) {
        // periodic images: cell-centred and face-centred along x / y
        const int ic = i < 0 ? i + nx : (i > nx - 1 ? i - nx : i);
        const int jc = j < 0 ? j + ny : (j > ny - 1 ? j - ny : j);
        const int if_ = i < 0 ? i + nx : (i > nx ? i - nx : i);
        const int jf = j < 0 ? j + ny : (j > ny ? j - ny : j);
        const long q = L.ix(i, j, k);
        const bool kval = k >= 0 && k <= nz - 1;
        const int depth = k < 0 ? -k : (k > nz - 1 ? k - nz + 1 : 0);   // vertical ghost row
        // ---- cons ghosts
        if (kval && (i < 0 || i > nx - 1 || j < 0 || j > ny - 1) && i >= -ngC && i <= nx - 1 + ngC &&
            j >= -ngC && j <= ny - 1 + ngC) {
            const long src = L.ix(ic, jc, k);
            for (int n = 0; n < nc; ++n) c[q + n * ne] = c[src + n * ne];
        }
        if (depth > 0) {
            const long src = L.ix(ic, jc, k < 0 ? 0 : nz - 1);
            for (int n = 0; n < nc; ++n) {
                const int rr = R.r[n];
                if (depth <= rr && i >= -rr && i <= nx - 1 + rr && j >= -rr && j <= ny - 1 + rr)
                    c[q + n * ne] = c[src + n * ne];
            }
        }
        // rho after the fills at a cell of valid k (any lateral index)
        auto rho_v = [&](int ii, int jj, int kk) {
            int a = ii < 0 ? ii + nx : (ii > nx - 1 ? ii - nx : ii);
            int b = jj < 0 ? jj + ny : (jj > ny - 1 ? jj - ny : jj);
            return c[L.ix(a, b, kk)];
        };
        // M2V of a valid face (ERF_MomentumToVelocity.cpp:24-96)
        auto u_of = [&](int fi, int jj, int kk) {
            long f = L.ix(fi, jj, kk);
            return xm[f] * 2.0 / (rho_v(fi, jj, kk) + rho_v(fi - 1, jj, kk));
        };
        auto v_of = [&](int ii, int fj, int kk) {
            long f = L.ix(ii, fj, kk);
            return ym[f] * 2.0 / (rho_v(ii, fj, kk) + rho_v(ii, fj - 1, kk));
        };
        auto w_of = [&](int ii, int jj, int fk) {
            long f = L.ix(ii, jj, fk);
            int kl = fk - 1 < 0 ? 0 : fk - 1, kh = fk > nz - 1 ? nz - 1 : fk;
            return zm[f] * 2.0 / (rho_v(ii, jj, kh) + rho_v(ii, jj, kl));
        };
        // ---- x-velocity: valid faces, FillBoundary layers, vertical ghost rows
        if (kval && i >= -ngV && i <= nx + ngV && j >= -ngV && j <= ny - 1 + ngV) {
            xv[q] = u_of(if_, jc, k);
        } else if (depth > 0 && depth <= ngv && i >= -ngv && i <= nx + ngv && j >= -ngv && j <= ny - 1 + ngv) {
            bool lo = k < 0;
            int t = lo ? xtl : xth;
            xv[q] = t == EXT_DIR ? (lo ? xvl : xvh) : u_of(if_, jc, lo ? 0 : nz - 1);
        }
        // ---- y-velocity
        if (kval && i >= -ngV && i <= nx - 1 + ngV && j >= -ngV && j <= ny + ngV) {
            yv[q] = v_of(ic, jf, k);
        } else if (depth > 0 && depth <= ngv && i >= -ngv && i <= nx - 1 + ngv && j >= -ngv && j <= ny + ngv) {
            bool lo = k < 0;
            int t = lo ? ytl : yth;
            yv[q] = t == EXT_DIR ? (lo ? yvl : yvh) : v_of(ic, jf, lo ? 0 : nz - 1);
        }
        // ---- z-velocity (k = 0 takes the zlo value on the valid box grown by ngv)
        if (k >= 0 && k <= nz && i >= -ngV && i <= nx - 1 + ngV && j >= -ngV && j <= ny - 1 + ngV) {
            const bool near = i >= -ngv && i <= nx - 1 + ngv && j >= -ngv && j <= ny - 1 + ngv;
            double w;
            if (terr && near && (k == 0 || (k == nz && zth == EXT_DIR))) {
                // terrain wall: WFromOmega with u, v as this launch leaves them
                // (the x / y branches above, evaluated again here)
                auto uf = [&](int ii, int jj, int kk) {
                    const int fi = ii < 0 ? ii + nx : (ii > nx ? ii - nx : ii);
                    const int cj = jj < 0 ? jj + ny : (jj > ny - 1 ? jj - ny : jj);
                    if (kk >= 0 && kk <= nz - 1) return u_of(fi, cj, kk);
                    const bool lo = kk < 0;
                    return (lo ? xtl : xth) == EXT_DIR ? (lo ? xvl : xvh) : u_of(fi, cj, lo ? 0 : nz - 1);
                };
                auto vf = [&](int ii, int jj, int kk) {
                    const int ci = ii < 0 ? ii + nx : (ii > nx - 1 ? ii - nx : ii);
                    const int fj = jj < 0 ? jj + ny : (jj > ny ? jj - ny : jj);
                    if (kk >= 0 && kk <= nz - 1) return v_of(ci, fj, kk);
                    const bool lo = kk < 0;
                    return (lo ? ytl : yth) == EXT_DIR ? (lo ? yvl : yvh) : v_of(ci, fj, lo ? 0 : nz - 1);
                };
                w = met::w_from_omega(i, j, k, k == 0 ? zvl : zvh, uf, vf, zt, dxi, dyi);
            } else if (near && k == 0) w = zvl;
            else if (near && k == nz && zth == EXT_DIR) w = zvh;
            else if (near && k == nz) w = (4.0 * w_of(ic, jc, nz - 1) - w_of(ic, jc, nz - 2)) / 3.0;
            else w = w_of(ic, jc, k);
            zv[q] = w;
        }
    });

    launch_box(grownB, s, [=] __device__(int i, int j, int k) {
        const int ic = i < 0 ? i + nx : (i > nx - 1 ? i - nx : i);
        const int jc = j < 0 ? j + ny : (j > ny - 1 ? j - ny : j);
        const int if_ = i < 0 ? i + nx : (i > nx ? i - nx : i);
        const int jf = j < 0 ? j + ny : (j > ny ? j - ny : j);
        const long q = L.ix(i, j, k);
        const bool kval = k >= 0 && k <= nz - 1;
        // VelocityToMomentum (ERF_VelocityToMomentum.cpp:24-114) at face index f
        auto xm_at = [&](long f) { return xv[f] * 0.5 * (c[f] + c[f - 1]); };
        auto ym_at = [&](long f) { return yv[f] * 0.5 * (c[f] + c[f - sy]); };
        auto zm_at = [&](long f) { return zv[f] * 0.5 * (c[f] + c[f - sz]); };
        // x-momentum: FillBoundary layers take the image's V2M
        if (kval && (i < 0 || i > nx || j < 0 || j > ny - 1) && i >= -ngV && i <= nx + ngV && j >= -ngV &&
            j <= ny - 1 + ngV) {
            xm[q] = xm_at(L.ix(if_, jc, k));
        } else if (i >= -1 && i <= nx + 1 && j >= -1 && j <= ny && k >= -1 && k <= nz) {
            xm[q] = xm_at(q);
        }
        if (kval && (i < 0 || i > nx - 1 || j < 0 || j > ny) && i >= -ngV && i <= nx - 1 + ngV && j >= -ngV &&
            j <= ny + ngV) {
            ym[q] = ym_at(L.ix(ic, jf, k));
        } else if (i >= -1 && i <= nx && j >= -1 && j <= ny + 1 && k >= -1 && k <= nz) {
            ym[q] = ym_at(q);
        }
        if (k >= 0 && k <= nz && (i < 0 || i > nx - 1 || j < 0 || j > ny - 1) && i >= -ngV && i <= nx - 1 + ngV &&
            j >= -ngV && j <= ny - 1 + ngV) {
            zm[q] = zm_at(L.ix(ic, jc, k));
        } else if (i >= -1 && i <= nx && j >= -1 && j <= ny && k >= 0 && k <= nz) {
            zm[q] = zm_at(q);
        }
    });
}

// ERF::FillIntermediatePatch, level 0 (ERF_FillIntermediatePatch.cpp:27-261).
void fill_intermediate_patch(const Params& P, const erfx_state& S, double* xv, double* yv, double* zv,
                             int ngc, int ngv, bool cons_only, int icomp, int ncomp, hipStream_t s) {
    if (!cons_only) momentum_to_velocity(P, xv, yv, zv, S.cons, S.xmom, S.ymom, S.zmom, s);
    fill_cons(P, S.cons, icomp, ncomp, ngc, s);
    if (!cons_only) {
        fill_xvel(P, xv, ngv, s);
        fill_yvel(P, yv, ngv, s);
        fill_zvel(P, zv, ngv, s, xv, yv);
        velocity_to_momentum(P, xv, yv, zv, 1, 0, S.cons, S.xmom, S.ymom, S.zmom, s);
    }
    fill_boundary_periodic(P, S.cons, P.cfg.ncons, P.ng_cons, -1, s);
    fill_boundary_periodic(P, S.xmom, 1, P.ng_vel, 0, s);
    fill_boundary_periodic(P, S.ymom, 1, P.ng_vel, 1, s);
    fill_boundary_periodic(P, S.zmom, 1, P.ng_vel, 2, s);
}

// apply_bcs (Source/TimeIntegration/ERF_TI_utils.H:53-150).
void apply_bcs(const Params& P, const erfx_state& S, double* xv, double* yv, double* zv, int ngc, int ngv,
               bool fast_only, bool synced, hipStream_t s) {
    if (!synced && ngc >= 1 && ngv >= 1 && fast_bcs_fusable(P)) {
        apply_bcs_periodic(P, S, xv, yv, zv, ngc, ngv, fast_only, s);
        return;
    }
    int scomp, ncomp, ngu;
    if (!synced) {
        ngu = ngc > ngv + 1 ? ngc : ngv + 1;
        fill_intermediate_patch(P, S, xv, yv, zv, ngu, 0, true, 0, 1, s);
    }
    if (synced) {
        scomp = 0;
        ncomp = fast_only ? 2 : P.cfg.ncons;
        ngu = ngc > ngv + 1 ? ngc : ngv + 1;
    } else {
        scomp = 1;
        ncomp = fast_only ? 1 : P.cfg.ncons - 1;
        ngu = ngc;
    }
    fill_intermediate_patch(P, S, xv, yv, zv, ngu, ngv, false, scomp, ncomp, s);
}

// ERF::FillPatch, level 0 at the old time (ERF_FillPatch.cpp:19-140).
void fill_patch(const Params& P, double* c, double* u, double* v, double* w, hipStream_t s) {
    fill_cons(P, c, 0, P.cfg.ncons, P.ng_cons, s);
    fill_xvel(P, u, P.ng_vel, s);
    fill_yvel(P, v, P.ng_vel, s);
    fill_zvel(P, w, P.ng_vel, s, u, v);
}

// cons_to_prim (ERF_TI_utils.H:4-45): prim on the box grown by ng, pi_stage
// on the box grown by one.  One launch covers both.
void cons_to_prim(const Params& P, const double* c, double* prim, double* pi_stage, int ng, hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int nprim = P.cfg.ncons - 1;
    const long ne = L.ne;
    const double rdOcp = P.rdOcp;
    Box b(-ng, -ng, -ng, nx - 1 + ng, ny - 1 + ng, nz - 1 + ng);
    launch_box(b, s, [=] __device__(int i, int j, int k) {
        long q = L.ix(i, j, k);
        double rho = c[q];
        double rho_theta = c[q + ne];
        prim[q] = rho_theta / rho;
        for (int n = 1; n < nprim; ++n) prim[q + n * ne] = c[q + (1 + n) * ne] / rho;
        if (i >= -1 && i <= nx && j >= -1 && j <= ny && k >= -1 && k <= nz)
            pi_stage[q] = eos_exner_of_rt(rho_theta, rdOcp);
    });
}

}  // namespace erfx

namespace erfx {

// Copy n doubles, 16 bytes per lane where both ends share their offset within
// 16 bytes (a field allocated with its rows shifted onto cache lines starts 8
// bytes into a 16-byte unit: its first double is then copied on its own), 8
// bytes per lane otherwise.  `first`/`stride` are the lane's index and the
// grid's width in lanes.
__device__ __forceinline__ void copy_span(double* dst, const double* src, long n, long first, long stride) {
    const unsigned long long as = reinterpret_cast<unsigned long long>(src);
    const unsigned long long ad = reinterpret_cast<unsigned long long>(dst);
    if (((as ^ ad) & 15ull) != 0) {
        for (long t = first; t < n; t += stride) dst[t] = src[t];
        return;
    }
    const long lead = (as & 15ull) ? 1 : 0;
    if (lead && first == 0 && n > 0) dst[0] = src[0];
    const long n2 = (n - lead) / 2;
    const double2* s2 = reinterpret_cast<const double2*>(src + lead);
    double2* d2 = reinterpret_cast<double2*>(dst + lead);
    for (long t = first; t < n2; t += stride) d2[t] = s2[t];
    const long tail = lead + 2 * n2;
    if (tail < n && first == 0) dst[tail] = src[tail];
}

__global__ void __launch_bounds__(256) copy_kernel(CopyList c) {
    const long stride = (long)gridDim.x * 256;
    for (int a = 0; a < c.count; ++a) copy_span(c.dst[a], c.src[a], c.n[a], (long)blockIdx.x * 256 + threadIdx.x, stride);
}

// batched form: each member's lists copied by its own range of work-groups
struct CopyItem {
    BHdr h;
    CopyList c;
};
__global__ void __launch_bounds__(256) copy_kernel_batch(const CopyItem* __restrict__ it,
                                                         const unsigned* __restrict__ starts, int n) {
    const CopyItem& m = batch_member(it, starts, n, blockIdx.x);
    const BHdr h = m.h;
    const CopyList& c = m.c;
    const long stride = (long)h.tiles * 256;
    const long first = (long)(blockIdx.x - h.t0) * 256 + threadIdx.x;
    for (int a = 0; a < c.count; ++a) copy_span(c.dst[a], c.src[a], c.n[a], first, stride);
}
static void copy_batch(const void* items, const unsigned* starts, int n, unsigned total, unsigned width,
                       hipStream_t s) {
    (void)width;
    hipLaunchKernelGGL(copy_kernel_batch, dim3(total), dim3(256), 0, s, static_cast<const CopyItem*>(items), starts,
                       n);
}

void copy_fields(const CopyList& c, hipStream_t s) {
    if (g_rec) {
        long nmax = 0;
        for (int a = 0; a < c.count; ++a) nmax = c.n[a] > nmax ? c.n[a] : nmax;
        const long want = (nmax / 2 + 255) / 256;
        CopyItem it{{{0, 0, 0}, {0, 0, 0}, 1u, 1u, 0u, (unsigned)(want < 1024 ? want : 1024)}, c};
        if (c.count == 0) it.h.tiles = 0;
        rec_push(&copy_batch, 256u, it);
        return;
    }
    if (c.count == 0) return;
    hipLaunchKernelGGL(copy_kernel, dim3(4096), dim3(256), 0, s, c);
}

}  // namespace erfx
