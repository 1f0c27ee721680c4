// erfx_terrain.h -- metric terms of a terrain-following grid (static terrain,
// unit map factors), evaluated on the node heights z_phys_nd
// (Source/Utils/ERF_TerrainMetrics.H:36-450).  Every function reads the
// node-height field through the box layout; the expressions keep the
// reference's operation order, so the results are those of the reference.
#pragma once

#include "erfx_common.h"

namespace erfx {

struct Zn {
    const double* z;
    Lay L;
    __device__ __forceinline__ double operator()(int i, int j, int k) const { return z[L.ix(i, j, k)]; }
};

// one field of the layout read at (i, j, k)
struct Fld {
    const double* p;
    Lay L;
    __device__ __forceinline__ double operator()(int i, int j, int k) const { return p[L.ix(i, j, k)]; }
};

namespace met {
__device__ __forceinline__ double h_xi_Iface(const Zn& z, int i, int j, int k, double dxi) {       // :105-116
    return 0.125 * dxi * (z(i + 1, j, k) + z(i + 1, j, k + 1) + z(i + 1, j + 1, k) + z(i + 1, j + 1, k + 1)
                          - z(i - 1, j, k) - z(i - 1, j, k + 1) - z(i - 1, j + 1, k) - z(i - 1, j + 1, k + 1));
}
__device__ __forceinline__ double h_zeta_Iface(const Zn& z, int i, int j, int k, double dzi) {     // :91-100
    return 0.5 * dzi * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k) - z(i, j + 1, k));
}
__device__ __forceinline__ double h_eta_Jface(const Zn& z, int i, int j, int k, double dyi) {      // :159-170
    return 0.125 * dyi * (z(i, j + 1, k) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k) + z(i + 1, j + 1, k + 1)
                          - z(i, j - 1, k) - z(i, j - 1, k + 1) - z(i + 1, j - 1, k) - z(i + 1, j - 1, k + 1));
}
__device__ __forceinline__ double h_zeta_Jface(const Zn& z, int i, int j, int k, double dzi) {     // :135-144
    return 0.5 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k) - z(i + 1, j, k));
}
__device__ __forceinline__ double h_zeta_Kface(const Zn& z, int i, int j, int k, double dzi) {     // :176-187
    return 0.125 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1)
                          - z(i, j, k - 1) - z(i + 1, j, k - 1) - z(i, j + 1, k - 1) - z(i + 1, j + 1, k - 1));
}
__device__ __forceinline__ double h_zeta_EdgeK(const Zn& z, int i, int j, int k, double dzi) {     // :225-233
    return dzi * (z(i, j, k + 1) - z(i, j, k));
}
__device__ __forceinline__ double h_zeta_EdgeJ(const Zn& z, int i, int j, int k, double dzi) {     // :268-277
    return 0.25 * dzi * (z(i, j, k + 1) + z(i, j + 1, k + 1) - z(i, j, k - 1) - z(i, j + 1, k - 1));
}
__device__ __forceinline__ double h_zeta_EdgeI(const Zn& z, int i, int j, int k, double dzi) {     // :307-316
    return 0.25 * dzi * (z(i, j, k + 1) + z(i + 1, j, k + 1) - z(i, j, k - 1) - z(i + 1, j, k - 1));
}
// cell-centre, k-face and edge-centre slopes for the stresses and the
// diffusive fluxes (:54-83, :192-219, :236-262, :279-303, :318-344)
__device__ __forceinline__ double h_xi_Cell(const Zn& z, int i, int j, int k, double dxi) {
    return 0.25 * dxi * (z(i + 1, j, k) + z(i + 1, j + 1, k) + z(i + 1, j, k + 1) + z(i + 1, j + 1, k + 1)
                         - z(i, j, k) - z(i, j + 1, k) - z(i, j, k + 1) - z(i, j + 1, k + 1));
}
__device__ __forceinline__ double h_eta_Cell(const Zn& z, int i, int j, int k, double dyi) {
    return 0.25 * dyi * (z(i, j + 1, k) + z(i + 1, j + 1, k) + z(i, j + 1, k + 1) + z(i + 1, j + 1, k + 1)
                         - z(i, j, k) - z(i + 1, j, k) - z(i, j, k + 1) - z(i + 1, j, k + 1));
}
__device__ __forceinline__ double h_xi_Kface(const Zn& z, int i, int j, int k, double dxi) {
    return 0.5 * dxi * (z(i + 1, j, k) + z(i + 1, j + 1, k) - z(i, j, k) - z(i, j + 1, k));
}
__device__ __forceinline__ double h_eta_Kface(const Zn& z, int i, int j, int k, double dyi) {
    return 0.5 * dyi * (z(i, j + 1, k) + z(i + 1, j + 1, k) - z(i, j, k) - z(i + 1, j, k));
}
__device__ __forceinline__ double h_xi_EdgeK(const Zn& z, int i, int j, int k, double dxi) {
    return 0.25 * dxi * (z(i + 1, j, k) + z(i + 1, j, k + 1) - z(i - 1, j, k) - z(i - 1, j, k + 1));
}
__device__ __forceinline__ double h_eta_EdgeK(const Zn& z, int i, int j, int k, double dyi) {
    return 0.25 * dyi * (z(i, j + 1, k) + z(i, j + 1, k + 1) - z(i, j - 1, k) - z(i, j - 1, k + 1));
}
__device__ __forceinline__ double h_xi_EdgeJ(const Zn& z, int i, int j, int k, double dxi) {
    return 0.25 * dxi * (z(i + 1, j + 1, k) + z(i + 1, j, k) - z(i - 1, j + 1, k) - z(i - 1, j, k));
}
__device__ __forceinline__ double h_eta_EdgeJ(const Zn& z, int i, int j, int k, double dyi) {
    return dyi * (z(i, j + 1, k) - z(i, j, k));
}
__device__ __forceinline__ double h_xi_EdgeI(const Zn& z, int i, int j, int k, double dxi) {
    return dxi * (z(i + 1, j, k) - z(i, j, k));
}
__device__ __forceinline__ double h_eta_EdgeI(const Zn& z, int i, int j, int k, double dyi) {
    return 0.25 * dyi * (z(i + 1, j + 1, k) + z(i, j + 1, k) - z(i + 1, j - 1, k) - z(i, j - 1, k));
}
// slopes of the z-face (i, j, k) of OmegaFromW / WFromOmega (:368-381)
__device__ __forceinline__ double zlo_xi(const Zn& z, int i, int j, int k, double dxi) {
    return 0.5 * dxi * (z(i + 1, j + 1, k) + z(i + 1, j, k) - z(i, j + 1, k) - z(i, j, k));
}
__device__ __forceinline__ double zlo_eta(const Zn& z, int i, int j, int k, double dyi) {
    return 0.5 * dyi * (z(i + 1, j + 1, k) + z(i, j + 1, k) - z(i + 1, j, k) - z(i, j, k));
}
// u and v averaged to the z-face, extrapolated at k = 0 (:383-388)
template <class U>
__device__ __forceinline__ double u_at_zface(const U& u, int i, int j, int k) {
    return (k == 0) ? 1.5 * (0.5 * (u(i, j, k) + u(i + 1, j, k))) - 0.5 * (0.5 * (u(i, j, k + 1) + u(i + 1, j, k + 1)))
                    : 0.25 * (u(i, j, k - 1) + u(i + 1, j, k - 1) + u(i, j, k) + u(i + 1, j, k));
}
template <class V>
__device__ __forceinline__ double v_at_zface(const V& v, int i, int j, int k) {
    return (k == 0) ? 1.5 * (0.5 * (v(i, j, k) + v(i, j + 1, k))) - 0.5 * (0.5 * (v(i, j, k + 1) + v(i, j + 1, k + 1)))
                    : 0.25 * (v(i, j, k - 1) + v(i, j + 1, k - 1) + v(i, j, k) + v(i, j + 1, k));
}
// The two slope terms of OmegaFromW / WFromOmega on z-face (i, j, k):
// px = zlo_xi * u and py = zlo_eta * v at the face, so that OmegaFromW is
// (w - px) - py and WFromOmega (omega + px) + py, the rounding of both
// functions below.
template <class U, class V>
__device__ __forceinline__ void zface_slope_terms(int i, int j, int k, const U& u, const V& v, const Zn& z, double dxi,
                                                  double dyi, double& px, double& py) {
    const double mx = zlo_xi(z, i, j, k, dxi);
    const double my = zlo_eta(z, i, j, k, dyi);
    px = mx * u_at_zface(u, i, j, k);
    py = my * v_at_zface(v, i, j, k);
}
// OmegaFromW (:359-392) and WFromOmega (:397-445); u, v: anything read as u(i, j, k)
template <class U, class V>
__device__ __forceinline__ double omega_from_w(int i, int j, int k, double w, const U& u, const V& v,
                                               const Zn& z, double dxi, double dyi) {
    double mx = zlo_xi(z, i, j, k, dxi);
    double my = zlo_eta(z, i, j, k, dyi);
    double uu = u_at_zface(u, i, j, k);
    double vv = v_at_zface(v, i, j, k);
    return w - mx * uu - my * vv;
}
template <class U, class V>
__device__ __forceinline__ double w_from_omega(int i, int j, int k, double omega, const U& u, const V& v,
                                               const Zn& z, double dxi, double dyi) {
    double uu = u_at_zface(u, i, j, k);
    double vv = v_at_zface(v, i, j, k);
    double mx = zlo_xi(z, i, j, k, dxi);
    double my = zlo_eta(z, i, j, k, dyi);
    return omega + mx * uu + my * vv;
}
}  // namespace met

}  // namespace erfx
