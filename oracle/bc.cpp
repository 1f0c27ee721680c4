// bc.cpp -- TEST INFRASTRUCTURE ONLY (oracle).  Ghost-cell fills and
// momentum/velocity conversion, restated from ERF's level-0 fill path.
#include "oracle.hpp"

namespace orc {

// ERF::init_bcs (Source/Initialization/ERF_init_bcs.cpp:300-500): one physical
// type per face -> a logical type per variable.
void Level::init(const erfx_config& c, const erfx_layout& lay) {
    cfg = c;
    L = lay;
    ncons = c.ncons;
    moist = c.moisture_type != ERFX_MOIST_NONE;
    n_qstate = moist ? 3 : 0;
    if (ncons != (moist ? ERFX_NCONS_MOIST : ERFX_NCONS_DRY)) throw std::runtime_error("ncons does not match moisture_type");
    if (c.use_terrain && c.les_type != ERFX_LES_NONE)
        throw std::runtime_error("an LES model on a terrain grid is not restated");
    ng_cons = lay.ng;
    ng_vel = lay.ng - 1;
    rdOcp = R_d / c.c_p;
    gravity = c.use_gravity ? CONST_GRAV : 0.0;
    dxInv[0] = 1.0 / c.dx; dxInv[1] = 1.0 / c.dy; dxInv[2] = 1.0 / c.dz;
    dz_min = c.dz;
    for (int d = 0; d < 3; ++d) periodic[d] = (c.bc[d] == ERFX_BC_PERIODIC);
    for (int ori = 0; ori < 6; ++ori) {
        int dir = ori % 3;
        int t = c.bc[ori];
        for (int v = 0; v < 3; ++v) {
            vel_extdir[v][ori] = 0.0;
            int k;
            switch (t) {
                case ERFX_BC_PERIODIC:     k = INT_DIR; break;
                case ERFX_BC_SYMMETRY:     k = (v == dir) ? REFLECT_ODD : REFLECT_EVEN; break;
                case ERFX_BC_OUTFLOW:      k = (v == dir) ? NEUMANN_INT : FOEXTRAP; break;
                // no-slip wall: ext_dir with the face's <face>.velocity, its
                // normal component forced to 0 (ERF_init_bcs.cpp:184-200)
                case ERFX_BC_NO_SLIP_WALL:
                    k = EXT_DIR;
                    vel_extdir[v][ori] = (v == dir) ? 0.0 : c.wall_velocity[ori][v];
                    break;
                case ERFX_BC_SLIP_WALL:    k = (v == dir) ? EXT_DIR : FOEXTRAP; break;
                // inflow: every velocity component ext_dir with the face's
                // values (ERF_init_bcs.cpp:117-120, 345-361)
                case ERFX_BC_INFLOW:       k = EXT_DIR; vel_extdir[v][ori] = c.inflow_velocity[ori][v]; break;
                default: throw std::runtime_error("bad bc");
            }
            bck[v][ori] = k;
        }
        int kc;
        switch (t) {
            case ERFX_BC_PERIODIC: kc = INT_DIR; break;
            case ERFX_BC_SYMMETRY: kc = REFLECT_EVEN; break;
            default:               kc = FOEXTRAP; break;   // outflow, slip and no-slip walls
        }
        for (int n = 0; n < ERFX_NCONS_MOIST; ++n) cons_extdir[n][ori] = 0.0;
        // inflow: every cons component ext_dir with density times the face's
        // value (m_bc_extdir_vals, ERF_init_bcs.cpp:125-163; 0 where not
        // given), or foextrap for all when the density is 0 (:544-580)
        if (t == ERFX_BC_INFLOW && c.inflow_density[ori] != 0.0) {
            const double rho_in = c.inflow_density[ori];
            kc = EXT_DIR;
            cons_extdir[ERFX_RHO][ori] = rho_in;
            cons_extdir[ERFX_RHOTHETA][ori] = rho_in * c.inflow_theta[ori];
            cons_extdir[ERFX_RHOSCALAR][ori] = rho_in * c.inflow_scalar[ori];
            if (c.moisture_type != ERFX_MOIST_NONE) {
                cons_extdir[ERFX_RHOQ1][ori] = rho_in * c.inflow_qv[ori];
                cons_extdir[ERFX_RHOQ2][ori] = rho_in * c.inflow_qc[ori];
            }
        }
        bck[3][ori] = kc;
    }

    cons.alloc(L, ncons); xvel.alloc(L, 1); yvel.alloc(L, 1); zvel.alloc(L, 1);
    r0.alloc(L, 1); p0.alloc(L, 1); pi0.alloc(L, 1);
    cons_old_vars.alloc(L, ncons); xvel_old.alloc(L, 1); yvel_old.alloc(L, 1); zvel_old.alloc(L, 1);
    S_old.alloc(L, ncons); S_new.alloc(L, ncons); S_sum.alloc(L, ncons);
    S_scratch.alloc(L, ncons); F_slow.alloc(L, ncons);
    prim.alloc(L, ncons - 1); pi_stage.alloc(L, 1); fast_coeffs.alloc(L, 5);
    omega.alloc(L, 1); eddy.alloc(L, EddyDiff::NumDiffs); hfx3.alloc(L, 1);
    tau11.alloc(L, 1); tau22.alloc(L, 1); tau33.alloc(L, 1);
    tau12.alloc(L, 1); tau13.alloc(L, 1); tau23.alloc(L, 1);
    tau21.alloc(L, 1); tau31.alloc(L, 1); tau32.alloc(L, 1);
    xmom_src.alloc(L, 1); ymom_src.alloc(L, 1); zmom_src.alloc(L, 1); cc_src.alloc(L, ncons);
    rain_accum.assign((size_t)L.sz, 0.0);
    if (c.use_terrain) {
        // flat node heights until set_terrain is given the real ones
        std::vector<double> zf((size_t)L.nelem);
        for (long q = 0; q < L.nelem; ++q) {
            long kk = q / L.sz - L.ng;
            zf[(size_t)q] = c.prob_lo[2] + (double)kk * c.dz;
        }
        set_terrain(zf.data());
    }
}

// FabArray::FillBoundary on a single periodic box: every ghost point whose
// periodic image lies in the valid region gets that value.  `nodal_dir` is the
// staggered direction (-1 for cell data).  Fills `ng` layers, all components.
void Level::fill_boundary_periodic(Arr a, int ncomp, int ng, int nodal_dir) {
    if (!periodic[0] && !periodic[1] && !periodic[2]) return;
    int n[3] = {L.nx, L.ny, L.nz};
    int vhi[3];
    for (int d = 0; d < 3; ++d) vhi[d] = n[d] - 1 + (d == nodal_dir ? 1 : 0);
    Box g(-ng, -ng, -ng, vhi[0] + ng, vhi[1] + ng, vhi[2] + ng);
    for (int c = 0; c < ncomp; ++c)
        loop(g, [&](int i, int j, int k) {
            int p[3] = {i, j, k};
            bool ghost = false;
            for (int d = 0; d < 3; ++d) if (p[d] < 0 || p[d] > vhi[d]) ghost = true;
            if (!ghost) return;
            int q[3] = {i, j, k};
            for (int d = 0; d < 3; ++d) {
                if (!periodic[d]) continue;
                if (q[d] < 0) q[d] += n[d];
                else if (q[d] > vhi[d]) q[d] -= n[d];
            }
            for (int d = 0; d < 3; ++d) if (q[d] < 0 || q[d] > vhi[d]) return;
            a(i, j, k, c) = a(q[0], q[1], q[2], c);
        });
}

// ERFPhysBCFunct_cons::operator() (ERF_PhysBCFunct.cpp:16-99) with do_fb = true,
// impose_lateral_cons_bcs (ERF_BoundaryConditions_cons.cpp:17-246) and
// impose_vertical_cons_bcs (:261-494) for the kinds reachable from the
// supported face types (foextrap, reflect_even; no Dirichlet/Neumann values).
void Level::fill_cons(Arr c, int icomp, int ncomp, int ng) {
    fill_boundary_periodic(c, ncons, ng_cons, -1);
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    // lateral: box = valid grown by ng in x,y (z valid only)
    if (!periodic[0]) {
        Box lo(-ng, -ng, 0, -1, ny - 1 + ng, nz - 1), hi(nx, -ng, 0, nx - 1 + ng, ny - 1 + ng, nz - 1);
        for (int n = icomp; n < icomp + ncomp; ++n) {
            int tl = bck[3][0], th = bck[3][3];
            // ext_dir (ERF_BoundaryConditions_cons.cpp:68-104): the face's value
            loop(lo, [&](int i, int j, int k) {
                if (tl == EXT_DIR) c(i, j, k, n) = cons_extdir[n][0];
                else if (tl == FOEXTRAP) c(i, j, k, n) = c(0, j, k, n);
                else if (tl == REFLECT_EVEN) c(i, j, k, n) = c(-1 - i, j, k, n);
            });
            loop(hi, [&](int i, int j, int k) {
                if (th == EXT_DIR) c(i, j, k, n) = cons_extdir[n][3];
                else if (th == FOEXTRAP) c(i, j, k, n) = c(nx - 1, j, k, n);
                else if (th == REFLECT_EVEN) c(i, j, k, n) = c(2 * (nx - 1) + 1 - i, j, k, n);
            });
        }
    }
    if (!periodic[1]) {
        Box lo(-ng, -ng, 0, nx - 1 + ng, -1, nz - 1), hi(-ng, ny, 0, nx - 1 + ng, ny - 1 + ng, nz - 1);
        for (int n = icomp; n < icomp + ncomp; ++n) {
            int tl = bck[3][1], th = bck[3][4];
            loop(lo, [&](int i, int j, int k) {
                if (tl == EXT_DIR) c(i, j, k, n) = cons_extdir[n][1];
                else if (tl == FOEXTRAP) c(i, j, k, n) = c(i, 0, k, n);
                else if (tl == REFLECT_EVEN) c(i, j, k, n) = c(i, -1 - j, k, n);
            });
            loop(hi, [&](int i, int j, int k) {
                if (th == EXT_DIR) c(i, j, k, n) = cons_extdir[n][4];
                else if (th == FOEXTRAP) c(i, j, k, n) = c(i, ny - 1, k, n);
                else if (th == REFLECT_EVEN) c(i, j, k, n) = c(i, 2 * (ny - 1) + 1 - j, k, n);
            });
        }
    }
    // vertical: box = valid grown by ng in all directions
    {
        Box lo(-ng, -ng, -ng, nx - 1 + ng, ny - 1 + ng, -1), hi(-ng, -ng, nz, nx - 1 + ng, ny - 1 + ng, nz - 1 + ng);
        int tl = bck[3][2], th = bck[3][5];
        if (tl != FOEXTRAP || th != FOEXTRAP) throw std::runtime_error("unsupported vertical cons bc");
        for (int n = icomp; n < icomp + ncomp; ++n) {
            loop(lo, [&](int i, int j, int k) { c(i, j, k, n) = c(i, j, 0, n); });
            loop(hi, [&](int i, int j, int k) { c(i, j, k, n) = c(i, j, nz - 1, n); });
        }
    }
}

// ERFPhysBCFunct_u::operator() (ERF_PhysBCFunct.cpp:101-186),
// impose_lateral_xvel_bcs / impose_vertical_xvel_bcs (ERF_BoundaryConditions_xvel.cpp:17-319).
void Level::fill_xvel(Arr u, int ng) {
    fill_boundary_periodic(u, 1, ng_vel, 0);
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    // xbx1: x-faces grown by ng, z clipped to the domain
    if (!periodic[0]) {
        int tl = bck[0][0], th = bck[0][3];
        double vl = vel_extdir[0][0], vh = vel_extdir[0][3];
        Box glo(-ng, -ng, 0, -1, ny - 1 + ng, nz - 1), flo(0, -ng, 0, 0, ny - 1 + ng, nz - 1);
        Box ghi(nx + 1, -ng, 0, nx + ng, ny - 1 + ng, nz - 1), fhi(nx, -ng, 0, nx, ny - 1 + ng, nz - 1);
        // the reference launches ghost and face fills together; neither reads
        // what the other writes for the supported kinds
        std::vector<double> face_lo, face_hi;
        loop(glo, [&](int i, int j, int k) {
            int iflip = -i;
            if (tl == EXT_DIR) u(i, j, k) = vl;
            else if (tl == FOEXTRAP) u(i, j, k) = u(0, j, k);
            else if (tl == REFLECT_EVEN) u(i, j, k) = u(iflip, j, k);
            else if (tl == REFLECT_ODD) u(i, j, k) = -u(iflip, j, k);
            else if (tl == NEUMANN_INT) u(i, j, k) = (4.0 * u(1, j, k) - u(2, j, k)) / 3.0;
        });
        loop(flo, [&](int i, int j, int k) {
            if (tl == EXT_DIR) u(i, j, k) = vl;
            else if (tl == NEUMANN_INT) u(i, j, k) = (4.0 * u(1, j, k) - u(2, j, k)) / 3.0;
        });
        loop(ghi, [&](int i, int j, int k) {
            int iflip = 2 * nx - i;
            if (th == EXT_DIR) u(i, j, k) = vh;
            else if (th == FOEXTRAP) u(i, j, k) = u(nx, j, k);
            else if (th == REFLECT_EVEN) u(i, j, k) = u(iflip, j, k);
            else if (th == REFLECT_ODD) u(i, j, k) = -u(iflip, j, k);
            else if (th == NEUMANN_INT) u(i, j, k) = (4.0 * u(nx - 1, j, k) - u(nx - 2, j, k)) / 3.0;
        });
        loop(fhi, [&](int i, int j, int k) {
            if (th == EXT_DIR) u(i, j, k) = vh;
            else if (th == NEUMANN_INT) u(i, j, k) = (4.0 * u(nx - 1, j, k) - u(nx - 2, j, k)) / 3.0;
        });
    }
    if (!periodic[1]) {
        int tl = bck[0][1], th = bck[0][4];
        double vl = vel_extdir[0][1], vh = vel_extdir[0][4];
        Box lo(-ng, -ng, 0, nx + ng, -1, nz - 1), hi(-ng, ny, 0, nx + ng, ny - 1 + ng, nz - 1);
        loop(lo, [&](int i, int j, int k) {
            int jflip = -1 - j;
            if (tl == EXT_DIR) u(i, j, k) = vl;
            else if (tl == FOEXTRAP) u(i, j, k) = u(i, 0, k);
            else if (tl == REFLECT_EVEN) u(i, j, k) = u(i, jflip, k);
            else if (tl == REFLECT_ODD) u(i, j, k) = -u(i, jflip, k);
        });
        loop(hi, [&](int i, int j, int k) {
            int jflip = 2 * (ny - 1) + 1 - j;
            if (th == EXT_DIR) u(i, j, k) = vh;
            else if (th == FOEXTRAP) u(i, j, k) = u(i, ny - 1, k);
            else if (th == REFLECT_EVEN) u(i, j, k) = u(i, jflip, k);
            else if (th == REFLECT_ODD) u(i, j, k) = -u(i, jflip, k);
        });
    }
    {   // vertical on xbx2 (x-faces grown by ng everywhere)
        int tl = bck[0][2], th = bck[0][5];
        double vl = vel_extdir[0][2], vh = vel_extdir[0][5];
        Box lo(-ng, -ng, -ng, nx + ng, ny - 1 + ng, -1), hi(-ng, -ng, nz, nx + ng, ny - 1 + ng, nz - 1 + ng);
        loop(lo, [&](int i, int j, int k) {
            int kflip = -1 - k;
            if (tl == EXT_DIR) u(i, j, k) = vl;
            else if (tl == FOEXTRAP) u(i, j, k) = u(i, j, 0);
            else if (tl == REFLECT_EVEN) u(i, j, k) = u(i, j, kflip);
            else if (tl == REFLECT_ODD) u(i, j, k) = -u(i, j, kflip);
        });
        loop(hi, [&](int i, int j, int k) {
            int kflip = 2 * (nz - 1) + 1 - k;
            if (th == EXT_DIR) u(i, j, k) = vh;
            else if (th == FOEXTRAP) u(i, j, k) = u(i, j, nz - 1);
            else if (th == REFLECT_EVEN) u(i, j, k) = u(i, j, kflip);
            else if (th == REFLECT_ODD) u(i, j, k) = -u(i, j, kflip);
        });
    }
}

// ERFPhysBCFunct_v::operator() (ERF_PhysBCFunct.cpp:188-269),
// ERF_BoundaryConditions_yvel.cpp:14-317.
void Level::fill_yvel(Arr v, int ng) {
    fill_boundary_periodic(v, 1, ng_vel, 1);
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    if (!periodic[0]) {
        int tl = bck[1][0], th = bck[1][3];
        double vl = vel_extdir[1][0], vh = vel_extdir[1][3];
        Box lo(-ng, -ng, 0, -1, ny + ng, nz - 1), hi(nx, -ng, 0, nx - 1 + ng, ny + ng, nz - 1);
        loop(lo, [&](int i, int j, int k) {
            int iflip = -1 - i;
            if (tl == EXT_DIR) v(i, j, k) = vl;
            else if (tl == FOEXTRAP) v(i, j, k) = v(0, j, k);
            else if (tl == REFLECT_EVEN) v(i, j, k) = v(iflip, j, k);
            else if (tl == REFLECT_ODD) v(i, j, k) = -v(iflip, j, k);
        });
        loop(hi, [&](int i, int j, int k) {
            int iflip = 2 * (nx - 1) + 1 - i;
            if (th == EXT_DIR) v(i, j, k) = vh;
            else if (th == FOEXTRAP) v(i, j, k) = v(nx - 1, j, k);
            else if (th == REFLECT_EVEN) v(i, j, k) = v(iflip, j, k);
            else if (th == REFLECT_ODD) v(i, j, k) = -v(iflip, j, k);
        });
    }
    if (!periodic[1]) {
        int tl = bck[1][1], th = bck[1][4];
        double vl = vel_extdir[1][1], vh = vel_extdir[1][4];
        Box glo(-ng, -ng, 0, nx - 1 + ng, -1, nz - 1), flo(-ng, 0, 0, nx - 1 + ng, 0, nz - 1);
        Box ghi(-ng, ny + 1, 0, nx - 1 + ng, ny + ng, nz - 1), fhi(-ng, ny, 0, nx - 1 + ng, ny, nz - 1);
        loop(glo, [&](int i, int j, int k) {
            int jflip = -j;
            if (tl == EXT_DIR) v(i, j, k) = vl;
            else if (tl == FOEXTRAP) v(i, j, k) = v(i, 0, k);
            else if (tl == REFLECT_EVEN) v(i, j, k) = v(i, jflip, k);
            else if (tl == REFLECT_ODD) v(i, j, k) = -v(i, jflip, k);
            else if (tl == NEUMANN_INT) v(i, j, k) = (4.0 * v(i, 1, k) - v(i, 2, k)) / 3.0;
        });
        loop(flo, [&](int i, int j, int k) {
            if (tl == EXT_DIR) v(i, j, k) = vl;
            else if (tl == NEUMANN_INT) v(i, j, k) = (4.0 * v(i, 1, k) - v(i, 2, k)) / 3.0;
        });
        loop(ghi, [&](int i, int j, int k) {
            int jflip = 2 * ny - j;
            if (th == EXT_DIR) v(i, j, k) = vh;
            else if (th == FOEXTRAP) v(i, j, k) = v(i, ny, k);
            else if (th == REFLECT_EVEN) v(i, j, k) = v(i, jflip, k);
            else if (th == REFLECT_ODD) v(i, j, k) = -v(i, jflip, k);
            else if (th == NEUMANN_INT) v(i, j, k) = (4.0 * v(i, ny - 1, k) - v(i, ny - 2, k)) / 3.0;
        });
        loop(fhi, [&](int i, int j, int k) {
            if (th == EXT_DIR) v(i, j, k) = vh;
            else if (th == NEUMANN_INT) v(i, j, k) = (4.0 * v(i, ny - 1, k) - v(i, ny - 2, k)) / 3.0;
        });
    }
    {
        int tl = bck[1][2], th = bck[1][5];
        double vl = vel_extdir[1][2], vh = vel_extdir[1][5];
        Box lo(-ng, -ng, -ng, nx - 1 + ng, ny + ng, -1), hi(-ng, -ng, nz, nx - 1 + ng, ny + ng, nz - 1 + ng);
        loop(lo, [&](int i, int j, int k) {
            int kflip = -1 - k;
            if (tl == EXT_DIR) v(i, j, k) = vl;
            else if (tl == FOEXTRAP) v(i, j, k) = v(i, j, 0);
            else if (tl == REFLECT_EVEN) v(i, j, k) = v(i, j, kflip);
            else if (tl == REFLECT_ODD) v(i, j, k) = -v(i, j, kflip);
        });
        loop(hi, [&](int i, int j, int k) {
            int kflip = 2 * (nz - 1) + 1 - k;
            if (th == EXT_DIR) v(i, j, k) = vh;
            else if (th == FOEXTRAP) v(i, j, k) = v(i, j, nz - 1);
            else if (th == REFLECT_EVEN) v(i, j, k) = v(i, j, kflip);
            else if (th == REFLECT_ODD) v(i, j, k) = -v(i, j, kflip);
        });
    }
}

// ERFPhysBCFunct_w::operator() (ERF_PhysBCFunct.cpp:271-367) and
// ERF_BoundaryConditions_zvel.cpp:16-268 (flat grid).  The box is clipped to
// the domain in z, so z-ghost faces are never filled; w at k = 0 always takes
// the zlo Dirichlet value.
void Level::fill_zvel(Arr w, int ng, Arr u, Arr v) {
    fill_boundary_periodic(w, 1, ng_vel, 2);
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    if (!periodic[0]) {
        int tl = bck[2][0], th = bck[2][3];
        double vl = vel_extdir[2][0], vh = vel_extdir[2][3];
        Box lo(-ng, -ng, 0, -1, ny - 1 + ng, nz), hi(nx, -ng, 0, nx - 1 + ng, ny - 1 + ng, nz);
        loop(lo, [&](int i, int j, int k) {
            int iflip = -1 - i;
            if (tl == EXT_DIR) w(i, j, k) = vl;
            else if (tl == FOEXTRAP) w(i, j, k) = w(0, j, k);
            else if (tl == REFLECT_EVEN) w(i, j, k) = w(iflip, j, k);
            else if (tl == REFLECT_ODD) w(i, j, k) = -w(iflip, j, k);
        });
        loop(hi, [&](int i, int j, int k) {
            int iflip = 2 * (nx - 1) + 1 - i;
            if (th == EXT_DIR) w(i, j, k) = vh;
            else if (th == FOEXTRAP) w(i, j, k) = w(nx - 1, j, k);
            else if (th == REFLECT_EVEN) w(i, j, k) = w(iflip, j, k);
            else if (th == REFLECT_ODD) w(i, j, k) = -w(iflip, j, k);
        });
    }
    if (!periodic[1]) {
        int tl = bck[2][1], th = bck[2][4];
        double vl = vel_extdir[2][1], vh = vel_extdir[2][4];
        Box lo(-ng, -ng, 0, nx - 1 + ng, -1, nz), hi(-ng, ny, 0, nx - 1 + ng, ny - 1 + ng, nz);
        loop(lo, [&](int i, int j, int k) {
            int jflip = -1 - j;
            if (tl == EXT_DIR) w(i, j, k) = vl;
            else if (tl == FOEXTRAP) w(i, j, k) = w(i, 0, k);
            else if (tl == REFLECT_EVEN) w(i, j, k) = w(i, jflip, k);
            else if (tl == REFLECT_ODD) w(i, j, k) = -w(i, jflip, k);
        });
        loop(hi, [&](int i, int j, int k) {
            int jflip = 2 * (ny - 1) + 1 - j;
            if (th == EXT_DIR) w(i, j, k) = vh;
            else if (th == FOEXTRAP) w(i, j, k) = w(i, ny - 1, k);
            else if (th == REFLECT_EVEN) w(i, j, k) = w(i, jflip, k);
            else if (th == REFLECT_ODD) w(i, j, k) = -w(i, jflip, k);
        });
    }
    {
        double vl = vel_extdir[2][2], vh = vel_extdir[2][5];
        Box lo(-ng, -ng, 0, nx - 1 + ng, ny - 1 + ng, 0), hi(-ng, -ng, nz, nx - 1 + ng, ny - 1 + ng, nz);
        // with static terrain the wall value is Omega's and w follows the
        // slope (ERF_BoundaryConditions_zvel.cpp:220-257)
        Arr zn = terrain ? A(z_nd) : Arr();
        if (terrain) loop(lo, [&](int i, int j, int k) { w(i, j, k) = met::w_from_omega(i, j, k, vl, u, v, zn, dxInv); });
        else loop(lo, [&](int i, int j, int k) { w(i, j, k) = vl; });
        int th = bck[2][5];
        if (th == EXT_DIR) {
            if (terrain) loop(hi, [&](int i, int j, int k) { w(i, j, k) = met::w_from_omega(i, j, k, vh, u, v, zn, dxInv); });
            else loop(hi, [&](int i, int j, int k) { w(i, j, k) = vh; });
        }
        else if (th == NEUMANN_INT)
            loop(hi, [&](int i, int j, int k) { w(i, j, k) = (4.0 * w(i, j, nz - 1) - w(i, j, nz - 2)) / 3.0; });
    }
}

// MomentumToVelocity (Source/Utils/ERF_MomentumToVelocity.cpp:24-96) on the
// valid box; the lateral ext_dir corrections apply only to cons Dirichlet
// faces, which the supported face types never produce.
void Level::momentum_to_velocity(Arr xv, Arr yv, Arr zv, Arr dens, Arr xm, Arr ym, Arr zm) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    loop(Box(0, 0, 0, nx, ny - 1, nz - 1), [&](int i, int j, int k) {
        xv(i, j, k) = xm(i, j, k) * 2.0 / (dens(i, j, k) + dens(i - 1, j, k));
    });
    loop(Box(0, 0, 0, nx - 1, ny, nz - 1), [&](int i, int j, int k) {
        yv(i, j, k) = ym(i, j, k) * 2.0 / (dens(i, j, k) + dens(i, j - 1, k));
    });
    loop(Box(0, 0, 0, nx - 1, ny - 1, nz), [&](int i, int j, int k) {
        zv(i, j, k) = zm(i, j, k) * 2.0 / (dens(i, j, k) + dens(i, j, k - 1));
    });
}

// VelocityToMomentum (Source/Utils/ERF_VelocityToMomentum.cpp:24-114): faces of
// the valid box grown by ngrow; w never outside [0, nz].
void Level::velocity_to_momentum(Arr xv, int ngu[3], Arr yv, int ngv[3], Arr zv, int ngw[3],
                                 Arr dens, Arr xm, Arr ym, Arr zm) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    loop(Box(-ngu[0], -ngu[1], -ngu[2], nx + ngu[0], ny - 1 + ngu[1], nz - 1 + ngu[2]), [&](int i, int j, int k) {
        xm(i, j, k) = xv(i, j, k) * 0.5 * (dens(i, j, k) + dens(i - 1, j, k));
    });
    loop(Box(-ngv[0], -ngv[1], -ngv[2], nx - 1 + ngv[0], ny + ngv[1], nz - 1 + ngv[2]), [&](int i, int j, int k) {
        ym(i, j, k) = yv(i, j, k) * 0.5 * (dens(i, j, k) + dens(i, j - 1, k));
    });
    int zlo = std::max(-ngw[2], 0), zhi = std::min(nz + ngw[2], nz);
    loop(Box(-ngw[0], -ngw[1], zlo, nx - 1 + ngw[0], ny - 1 + ngw[1], zhi), [&](int i, int j, int k) {
        zm(i, j, k) = zv(i, j, k) * 0.5 * (dens(i, j, k) + dens(i, j, k - 1));
    });
}

// ERF::FillIntermediatePatch, level 0 (ERF_FillIntermediatePatch.cpp:27-261).
void Level::fill_intermediate_patch(StateA S, Arr xv, Arr yv, Arr zv, int ngc, int ngv,
                                    bool cons_only, int icomp, int ncomp) {
    if (!cons_only) momentum_to_velocity(xv, yv, zv, S.cons, S.xmom, S.ymom, S.zmom);
    fill_cons(S.cons, icomp, ncomp, ngc);
    if (!cons_only) {
        fill_xvel(xv, ngv);
        fill_yvel(yv, ngv);
        fill_zvel(zv, ngv, xv, yv);
        int g1[3] = {1, 1, 1}, g0[3] = {1, 1, 0};
        velocity_to_momentum(xv, g1, yv, g1, zv, g0, S.cons, S.xmom, S.ymom, S.zmom);
    }
    fill_boundary_periodic(S.cons, ncons, ng_cons, -1);
    fill_boundary_periodic(S.xmom, 1, ng_vel, 0);
    fill_boundary_periodic(S.ymom, 1, ng_vel, 1);
    fill_boundary_periodic(S.zmom, 1, ng_vel, 2);
}

// apply_bcs (Source/TimeIntegration/ERF_TI_utils.H:53-150).
void Level::apply_bcs(StateA S, Arr xv, Arr yv, Arr zv, int ngc, int ngv, bool fast_only,
                      bool synced) {
    int scomp, ncomp, ngu;
    if (!synced) {
        ngu = std::max(ngc, ngv + 1);
        fill_intermediate_patch(S, xv, yv, zv, ngu, 0, true, 0, 1);
    }
    if (synced) {
        scomp = 0;
        ncomp = fast_only ? 2 : ncons;
        ngu = std::max(ngc, ngv + 1);
    } else {
        scomp = 1;
        ncomp = fast_only ? 1 : ncons - 1;
        ngu = ngc;
    }
    fill_intermediate_patch(S, xv, yv, zv, ngu, ngv, false, scomp, ncomp);
}

// ERF::FillPatch, level 0, at the old time (ERF_FillPatch.cpp:19-140): periodic
// and physical fills of cons and velocity over their full ghost widths.
void Level::fill_patch(Arr c, Arr u, Arr v, Arr w) {
    fill_cons(c, 0, ncons, ng_cons);
    fill_xvel(u, ng_vel);
    fill_yvel(v, ng_vel);
    fill_zvel(w, ng_vel, u, v);
}

// cons_to_prim (ERF_TI_utils.H:4-45)
void Level::cons_to_prim(Arr c, int ng) {
    Arr pr = A(prim), pis = A(pi_stage);
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    int nprim = ncons - 1;
    loop(Box(-ng, -ng, -ng, nx - 1 + ng, ny - 1 + ng, nz - 1 + ng), [&](int i, int j, int k) {
        double rho = c(i, j, k, 0);
        double rho_theta = c(i, j, k, 1);
        pr(i, j, k, 0) = rho_theta / rho;
        for (int n = 1; n < nprim; ++n) pr(i, j, k, n) = c(i, j, k, 1 + n) / rho;
    });
    loop(Box(-1, -1, -1, nx, ny, nz), [&](int i, int j, int k) {
        pis(i, j, k) = getExnergivenRTh(c(i, j, k, 1), rdOcp);
    });
}

}  // namespace orc
