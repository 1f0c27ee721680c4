// erfx_slow.hip -- slow right-hand side of one RK stage (flat, dry):
// Smagorinsky eddy viscosity, stress, momentum sources, erf_slow_rhs_pre and
// erf_slow_rhs_post as gfx950 kernels.
//
// Fluxes are evaluated on the fly: a cell kernel computes both faces of its
// flux divergence, so a face flux is evaluated by the two cells that share it
// with the identical expression (bitwise the same value the reference stores
// in its flux arrays).  Face kernels fuse advection, diffusion, pressure
// gradient and sources of one momentum component.
#include "erfx_common.h"
#include "erfx_internal.h"
#include "erfx_terrain.h"
#include <type_traits>

namespace erfx {

// ---- face interpolation (ERF_Interpolation_UPW.H) --------------------------
// Value on the lo face of point s along a line; sm3..sp2 are the values at
// offsets -3..+2.  CENTERED2 :9-87, UPWIND3 :92-201, CENTERED4 :207-295,
// UPWIND5 :300-395, CENTERED6 :400-480.
// WENO family for scalars (Source/Utils/ERF_Interpolation_WENO.H: WENO3 :9-127,
// WENO5 :132-262; ERF_Interpolation_WENO_Z.H: WENO_Z3 :9-126, WENO_MZQ3
// :130-252, WENO_Z5 :256-389).  The stencil is passed from the upwind side.
namespace weno {
constexpr double eps = 1.0e-6, tol = 1.0e-12;
__device__ __forceinline__ double w3(double sm1, double s, double sp1) {
    const double g1 = (1.0 / 3.0), g2 = (2.0 / 3.0);
    double b1 = (s - sm1) * (s - sm1);
    double b2 = (sp1 - s) * (sp1 - s);
    double w1 = g1 / ((eps + b1) * (eps + b1));
    double w2 = g2 / ((eps + b2) * (eps + b2));
    double wsum = w1 + w2;
    double v1 = -sm1 + 3.0 * s;
    double v2 = s + sp1;
    return ((w1 * v1 + w2 * v2) / (2.0 * wsum));
}
__device__ __forceinline__ double z3(double sm1, double s, double sp1) {
    const double g1 = (1.0 / 3.0), g2 = (2.0 / 3.0);
    double b1 = (s - sm1) * (s - sm1);
    double b2 = (sp1 - s) * (sp1 - s);
    double t5 = fabs(b2 - b1);
    double w1 = g1 * (1.0 + (t5 * t5) / ((eps + b1) * (eps + b1)));
    double w2 = g2 * (1.0 + (t5 * t5) / ((eps + b2) * (eps + b2)));
    double wsum = w1 + w2;
    double v1 = -sm1 + 3.0 * s;
    double v2 = s + sp1;
    return ((w1 * v1 + w2 * v2) / (2.0 * wsum));
}
__device__ __forceinline__ double mzq3(double sm1, double s, double sp1) {
    const double g1 = (1.0 / 3.0), g2 = (1.0 / 3.0), g3 = (1.0 / 3.0);
    double b1 = (s - sm1) * (s - sm1);
    double b2 = (sp1 - s) * (sp1 - s);
    double b3 = ((13.0 / 12.0) * ((sm1 - 2.0 * s + sp1) * (sm1 - 2.0 * s + sp1))) + (((sp1 - sm1) * (sp1 - sm1)) / 4.0);
    double t5 = (fabs(b3 - b1) + fabs(b3 - b2)) / 32.0;
    double a1 = g1 * (1.0 + (t5 * t5) / ((eps + b1) * (eps + b1)));
    double a2 = g2 * (1.0 + (t5 * t5) / ((eps + b2) * (eps + b2)));
    double a3 = g3 * (1.0 + (t5 * t5) / ((eps + b3) * (eps + b3)));
    double asum = a1 + a2 + a3;
    double w1 = a1 / asum;
    double w2 = a2 / asum;
    double w3_ = a3 / asum;
    double v1 = (-sm1 + 3.0 * s) / 2.0;
    double v2 = (s + sp1) / 2.0;
    double v3 = (-sm1 + 5.0 * s + 2.0 * sp1) / 6.0;
    return ((w3_ / g3) * (v3 - g1 * v1 - g2 * v2) + w1 * v1 + w2 * v2);
}
__device__ __forceinline__ double w5(double sm2, double sm1, double s, double sp1, double sp2, bool z) {
    const double c1 = (13.0 / 12.0);
    const double g1 = (1.0 / 10.0), g2 = (3.0 / 5.0), g3 = (3.0 / 10.0);
    double b1 = c1 * (sm2 - 2.0 * sm1 + s) * (sm2 - 2.0 * sm1 + s) +
                0.25 * (sm2 - 4.0 * sm1 + 3.0 * s) * (sm2 - 4.0 * sm1 + 3.0 * s);
    double b2 = c1 * (sm1 - 2.0 * s + sp1) * (sm1 - 2.0 * s + sp1) + 0.25 * (sm1 - sp1) * (sm1 - sp1);
    double b3 = c1 * (s - 2.0 * sp1 + sp2) * (s - 2.0 * sp1 + sp2) +
                0.25 * (3.0 * s - 4.0 * sp1 + sp2) * (3.0 * s - 4.0 * sp1 + sp2);
    double w1, w2, w3_;
    if (z) {
        double t5 = fabs(b3 - b1);
        w1 = g1 * (1.0 + (t5 * t5) / ((eps + b1) * (eps + b1)));
        w2 = g2 * (1.0 + (t5 * t5) / ((eps + b2) * (eps + b2)));
        w3_ = g3 * (1.0 + (t5 * t5) / ((eps + b3) * (eps + b3)));
    } else {
        w1 = g1 / ((eps + b1) * (eps + b1));
        w2 = g2 / ((eps + b2) * (eps + b2));
        w3_ = g3 / ((eps + b3) * (eps + b3));
    }
    double wsum = w1 + w2 + w3_;
    double v1 = 2.0 * sm2 - 7.0 * sm1 + 11.0 * s;
    double v2 = -sm1 + 5.0 * s + 2.0 * sp1;
    double v3 = 2.0 * s + 5.0 * sp1 - sp2;
    return ((w1 * v1 + w2 * v2 + w3_ * v3) / (6.0 * wsum));
}
__device__ __forceinline__ double face(int t, double sp2, double sp1, double s, double sm1, double sm2, double sm3,
                                       double upw) {
    if (!(upw > tol) && !(upw < -tol)) return 0.5 * (s + sm1);
    // The stencil seen from the upwind side, far upwind point first: the
    // operands are selected per lane and the weights evaluated once, so lanes
    // of one wave with opposite flow directions do not run both evaluations
    // (the result is the same call on the same operands as the two-branch form).
    const bool pos = upw > tol;
    const double a = pos ? sm3 : sp2, b = pos ? sm2 : sp1, c = pos ? sm1 : s, d = pos ? s : sm1, e = pos ? sp1 : sm2;
    if (t == ERFX_ADV_WENO3) return w3(b, c, d);
    if (t == ERFX_ADV_WENOZ3) return z3(b, c, d);
    if (t == ERFX_ADV_WENOMZQ3) return mzq3(b, c, d);
    return w5(a, b, c, d, e, t == ERFX_ADV_WENOZ5);
}
}  // namespace weno

__host__ __device__ __forceinline__ bool is_weno(int t) { return t >= ERFX_ADV_WENO3 && t <= ERFX_ADV_WENOZ5; }

__device__ __forceinline__ double interp_line(int t, double sp2, double sp1, double s, double sm1, double sm2,
                                              double sm3, double upw, double upw_frac) {
    if (t == ERFX_ADV_CENTERED2) {
        double a1 = (s + sm1);
        return (0.5 * a1);
    }
    if (t == ERFX_ADV_UPWIND3) {
        if (upw != 0.) upw = (upw > 0) ? 1. : -1.;
        upw *= upw_frac;
        double a1 = (s + sm1), d1 = (s - sm1), a2 = (sp1 + sm2), d2 = (sp1 - sm2);
        const double g1 = (7.0 / 12.0), g2 = (1.0 / 12.0);
        return (g1 * a1 - g2 * a2 + upw * g2 * (d2 - 3.0 * d1));
    }
    if (t == ERFX_ADV_CENTERED4) {
        double a1 = (s + sm1), a2 = (sp1 + sm2);
        const double g1 = (7.0 / 12.0), g2 = (1.0 / 12.0);
        return (g1 * a1 - g2 * a2);
    }
    if (t == ERFX_ADV_UPWIND5) {
        if (upw != 0.) upw = (upw > 0) ? 1. : -1.;
        upw *= upw_frac;
        double a1 = (s + sm1), a2 = (sp1 + sm2), a3 = (sp2 + sm3);
        double d1 = (s - sm1), d2 = (sp1 - sm2), d3 = (sp2 - sm3);
        const double g1 = (37.0 / 60.0), g2 = (2.0 / 15.0), g3 = (1.0 / 60.0);
        return (g1 * a1 - g2 * a2 + g3 * a3 - upw * g3 * (d3 - 5.0 * d2 + 10.0 * d1));
    }
    // CENTERED6
    double a1 = (s + sm1), a2 = (sp1 + sm2), a3 = (sp2 + sm3);
    const double g1 = (37.0 / 60.0), g2 = (2.0 / 15.0), g3 = (1.0 / 60.0);
    return (g1 * a1 - g2 * a2 + g3 * a3);
}

__device__ __forceinline__ bool wide_stencil(int t) { return t == ERFX_ADV_UPWIND5 || t == ERFX_ADV_CENTERED6; }

// interpolate field f along stride st at index q (lo face of q)
__device__ __forceinline__ double interp_at(int t, const double* f, long q, long st, double upw, double fr) {
    if (t == ERFX_ADV_CENTERED2) return interp_line(t, 0, 0, f[q], f[q - st], 0, 0, upw, fr);
    double sp2 = wide_stencil(t) ? f[q + 2 * st] : 0.;
    double sm3 = wide_stencil(t) ? f[q - 3 * st] : 0.;
    return interp_line(t, sp2, f[q + st], f[q], f[q - st], f[q - 2 * st], sm3, upw, fr);
}

// scalar interpolation: the WENO family or the centred/upwind family
__device__ __forceinline__ double interp_scalar_at(int t, const double* f, long q, long st, double upw, double fr) {
    if (is_weno(t)) {
        double sp2 = f[q + 2 * st], sm3 = f[q - 3 * st];
        return weno::face(t, sp2, f[q + st], f[q], f[q - st], f[q - 2 * st], sm3, upw);
    }
    return interp_at(t, f, q, st, upw, fr);
}

// ---- mixed precision: face values in FP32 on perturbations -------------------
// The same stencils and WENO weights in FP32, applied to the deviations of the
// stencil values from an FP64 reference c at the face (the mean of its two
// cells); the face value is c plus the FP32 result.  The interpolants are
// consistent (weights summing to one) and the smoothness indicators depend on
// differences only, so the shift changes nothing in exact arithmetic; the
// FP32 error scales with the local variation, not with the value itself.
namespace wenof {
constexpr float eps = 1.0e-6f;
// the hardware reciprocal (v_rcp_f32, 1 ulp): the FP32 weights need no
// correctly rounded division, and this form is a few instructions instead of
// the dozen of an IEEE division
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float w3(float sm1, float s, float sp1) {
    const float g1 = (1.0f / 3.0f), g2 = (2.0f / 3.0f);
    float b1 = (s - sm1) * (s - sm1);
    float b2 = (sp1 - s) * (sp1 - s);
    float w1 = g1 * rcp((eps + b1) * (eps + b1));
    float w2 = g2 * rcp((eps + b2) * (eps + b2));
    return ((w1 * (-sm1 + 3.0f * s) + w2 * (s + sp1)) * rcp(2.0f * (w1 + w2)));
}
__device__ __forceinline__ float z3(float sm1, float s, float sp1) {
    const float g1 = (1.0f / 3.0f), g2 = (2.0f / 3.0f);
    float b1 = (s - sm1) * (s - sm1);
    float b2 = (sp1 - s) * (sp1 - s);
    float t5 = fabsf(b2 - b1);
    float w1 = g1 * (1.0f + (t5 * t5) * rcp((eps + b1) * (eps + b1)));
    float w2 = g2 * (1.0f + (t5 * t5) * rcp((eps + b2) * (eps + b2)));
    return ((w1 * (-sm1 + 3.0f * s) + w2 * (s + sp1)) * rcp(2.0f * (w1 + w2)));
}
__device__ __forceinline__ float mzq3(float sm1, float s, float sp1) {
    const float g = (1.0f / 3.0f);
    float b1 = (s - sm1) * (s - sm1);
    float b2 = (sp1 - s) * (sp1 - s);
    float b3 = ((13.0f / 12.0f) * ((sm1 - 2.0f * s + sp1) * (sm1 - 2.0f * s + sp1))) + (((sp1 - sm1) * (sp1 - sm1)) / 4.0f);
    float t5 = (fabsf(b3 - b1) + fabsf(b3 - b2)) / 32.0f;
    float a1 = g * (1.0f + (t5 * t5) * rcp((eps + b1) * (eps + b1)));
    float a2 = g * (1.0f + (t5 * t5) * rcp((eps + b2) * (eps + b2)));
    float a3 = g * (1.0f + (t5 * t5) * rcp((eps + b3) * (eps + b3)));
    float asum = a1 + a2 + a3;
    float v1 = (-sm1 + 3.0f * s) / 2.0f, v2 = (s + sp1) / 2.0f, v3 = (-sm1 + 5.0f * s + 2.0f * sp1) / 6.0f;
    return ((a3 / asum / g) * (v3 - g * v1 - g * v2) + (a1 / asum) * v1 + (a2 / asum) * v2);
}
__device__ __forceinline__ float w5(float sm2, float sm1, float s, float sp1, float sp2, bool z) {
    const float c1 = (13.0f / 12.0f);
    const float g1 = (1.0f / 10.0f), g2 = (3.0f / 5.0f), g3 = (3.0f / 10.0f);
    float b1 = c1 * (sm2 - 2.0f * sm1 + s) * (sm2 - 2.0f * sm1 + s) +
               0.25f * (sm2 - 4.0f * sm1 + 3.0f * s) * (sm2 - 4.0f * sm1 + 3.0f * s);
    float b2 = c1 * (sm1 - 2.0f * s + sp1) * (sm1 - 2.0f * s + sp1) + 0.25f * (sm1 - sp1) * (sm1 - sp1);
    float b3 = c1 * (s - 2.0f * sp1 + sp2) * (s - 2.0f * sp1 + sp2) +
               0.25f * (3.0f * s - 4.0f * sp1 + sp2) * (3.0f * s - 4.0f * sp1 + sp2);
    float w1, w2, w3_;
    if (z) {
        float t5 = fabsf(b3 - b1);
        w1 = g1 * (1.0f + (t5 * t5) * rcp((eps + b1) * (eps + b1)));
        w2 = g2 * (1.0f + (t5 * t5) * rcp((eps + b2) * (eps + b2)));
        w3_ = g3 * (1.0f + (t5 * t5) * rcp((eps + b3) * (eps + b3)));
    } else {
        w1 = g1 * rcp((eps + b1) * (eps + b1));
        w2 = g2 * rcp((eps + b2) * (eps + b2));
        w3_ = g3 * rcp((eps + b3) * (eps + b3));
    }
    float v1 = 2.0f * sm2 - 7.0f * sm1 + 11.0f * s;
    float v2 = -sm1 + 5.0f * s + 2.0f * sp1;
    float v3 = 2.0f * s + 5.0f * sp1 - sp2;
    return ((w1 * v1 + w2 * v2 + w3_ * v3) * rcp(6.0f * (w1 + w2 + w3_)));
}
__device__ __forceinline__ float face(int t, float sp2, float sp1, float s, float sm1, float sm2, float sm3, double upw) {
    if (!(upw > weno::tol) && !(upw < -weno::tol)) return 0.5f * (s + sm1);
    // operands selected per lane, weights evaluated once (as in weno::face)
    const bool pos = upw > weno::tol;
    const float a = pos ? sm3 : sp2, b = pos ? sm2 : sp1, c = pos ? sm1 : s, d = pos ? s : sm1, e = pos ? sp1 : sm2;
    if (t == ERFX_ADV_WENO3) return w3(b, c, d);
    if (t == ERFX_ADV_WENOZ3) return z3(b, c, d);
    if (t == ERFX_ADV_WENOMZQ3) return mzq3(b, c, d);
    return w5(a, b, c, d, e, t == ERFX_ADV_WENOZ5);
}
}  // namespace wenof

__device__ __forceinline__ float interp_line_f(int t, float sp2, float sp1, float s, float sm1, float sm2, float sm3,
                                               double upw, double upw_frac) {
    if (t == ERFX_ADV_CENTERED2) return 0.5f * (s + sm1);
    float u = (upw != 0.) ? ((upw > 0) ? 1.0f : -1.0f) : 0.0f;
    u *= (float)upw_frac;
    if (t == ERFX_ADV_UPWIND3)
        return (7.0f / 12.0f) * (s + sm1) - (1.0f / 12.0f) * (sp1 + sm2) + u * (1.0f / 12.0f) * ((sp1 - sm2) - 3.0f * (s - sm1));
    if (t == ERFX_ADV_CENTERED4) return (7.0f / 12.0f) * (s + sm1) - (1.0f / 12.0f) * (sp1 + sm2);
    const float c5 = (37.0f / 60.0f) * (s + sm1) - (2.0f / 15.0f) * (sp1 + sm2) + (1.0f / 60.0f) * (sp2 + sm3);
    if (t == ERFX_ADV_UPWIND5)
        return c5 - u * (1.0f / 60.0f) * ((sp2 - sm3) - 5.0f * (sp1 - sm2) + 10.0f * (s - sm1));
    return c5;   // CENTERED6
}

// face value of f on the lo face of q along stride st, mixed precision
__device__ __forceinline__ double interp_mixed(int t, const double* f, long q, long st, double upw, double fr) {
    const double c = 0.5 * (f[q] + f[q - st]);
    if (t == ERFX_ADV_CENTERED2) return c;
    const bool wide = t == ERFX_ADV_WENO5 || t == ERFX_ADV_WENOZ5 || t == ERFX_ADV_WENO3 || t == ERFX_ADV_WENOZ3 ||
                      t == ERFX_ADV_WENOMZQ3 || wide_stencil(t);
    const float sp1 = (float)(f[q + st] - c), s = (float)(f[q] - c), sm1 = (float)(f[q - st] - c);
    const float sm2 = (float)(f[q - 2 * st] - c);
    const float sp2 = wide ? (float)(f[q + 2 * st] - c) : 0.f, sm3 = wide ? (float)(f[q - 3 * st] - c) : 0.f;
    const float v = is_weno(t) ? wenof::face(t, sp2, sp1, s, sm1, sm2, sm3, upw)
                               : interp_line_f(t, sp2, sp1, s, sm1, sm2, sm3, upw, fr);
    return c + (double)v;
}
// the FP64 or the mixed face value
__device__ __forceinline__ double interp_pick(bool mixed, int t, const double* f, long q, long st, double upw, double fr) {
    return mixed ? interp_mixed(t, f, q, st, upw, fr) : interp_at(t, f, q, st, upw, fr);
}
__device__ __forceinline__ double interp_scalar_pick(bool mixed, int t, const double* f, long q, long st, double upw,
                                                     double fr) {
    return mixed ? interp_mixed(t, f, q, st, upw, fr) : interp_scalar_at(t, f, q, st, upw, fr);
}
// The same face value from the six stencil values at offsets -3..+2 (sm3..sp2)
// already at hand (staged in LDS or kept in registers): the arithmetic of
// interp_scalar_pick, term for term.  Stencils narrower than six points do not
// read the outer values.
template <bool mixed>
__device__ __forceinline__ double interp_scalar_vals(int t, double sp2, double sp1, double s, double sm1, double sm2,
                                                     double sm3, double upw, double fr) {
    if (!mixed) {
        if (is_weno(t)) return weno::face(t, sp2, sp1, s, sm1, sm2, sm3, upw);
        return interp_line(t, sp2, sp1, s, sm1, sm2, sm3, upw, fr);
    }
    const double c = 0.5 * (s + sm1);
    if (t == ERFX_ADV_CENTERED2) return c;
    const bool wide = is_weno(t) || wide_stencil(t);
    const float fp1 = (float)(sp1 - c), f0 = (float)(s - c), fm1 = (float)(sm1 - c), fm2 = (float)(sm2 - c);
    const float fp2 = wide ? (float)(sp2 - c) : 0.f, fm3 = wide ? (float)(sm3 - c) : 0.f;
    const float v = is_weno(t) ? wenof::face(t, fp2, fp1, f0, fm1, fm2, fm3, upw)
                               : interp_line_f(t, fp2, fp1, f0, fm1, fm2, fm3, upw, fr);
    return c + (double)v;
}
// The same with the type fixed at compile time: TT < 0 takes t at run time;
// TT = U3 or U5 folds the type's branches; TT = WENO5 or WENO-Z5 is
// straight-line code (the upwind operands selected
// per lane, the weights always evaluated, the centred value selected where the
// face velocity is within the tolerance of zero), so the x-, y- and z-faces of
// a level can be evaluated interleaved.  The value is the same expression on
// the same operands as the run-time form.
template <bool mixed, int TT>
__device__ __forceinline__ double interp_scalar_vals_t(int t, double sp2, double sp1, double s, double sm1,
                                                       double sm2, double sm3, double upw, double fr) {
    if constexpr (TT < 0) {
        return interp_scalar_vals<mixed>(t, sp2, sp1, s, sm1, sm2, sm3, upw, fr);
    } else if constexpr (TT == ERFX_ADV_UPWIND3 || TT == ERFX_ADV_UPWIND5) {
        // linear stencils: the run-time form with the type a constant
        return interp_scalar_vals<mixed>(TT, sp2, sp1, s, sm1, sm2, sm3, upw, fr);
    } else {
        static_assert(TT == ERFX_ADV_WENO5 || TT == ERFX_ADV_WENOZ5, "fixed types: U3, U5, WENO5, WENO-Z5");
        const bool pos = upw > weno::tol;
        const bool upwinded = upw > weno::tol || upw < -weno::tol;
        if constexpr (!mixed) {
            const double a = pos ? sm3 : sp2, b = pos ? sm2 : sp1, c = pos ? sm1 : s, d = pos ? s : sm1;
            const double e = pos ? sp1 : sm2;
            const double r = weno::w5(a, b, c, d, e, TT == ERFX_ADV_WENOZ5);
            return upwinded ? r : 0.5 * (s + sm1);
        } else {
            const double c0 = 0.5 * (s + sm1);
            const float fp1 = (float)(sp1 - c0), f0 = (float)(s - c0), fm1 = (float)(sm1 - c0);
            const float fm2 = (float)(sm2 - c0), fp2 = (float)(sp2 - c0), fm3 = (float)(sm3 - c0);
            const float a = pos ? fm3 : fp2, b = pos ? fm2 : fp1, c = pos ? fm1 : f0, d = pos ? f0 : fm1;
            const float e = pos ? fp1 : fm2;
            const float r = wenof::w5(a, b, c, d, e, TT == ERFX_ADV_WENOZ5);
            const float v = upwinded ? r : 0.5f * (f0 + fm1);
            return c0 + (double)v;
        }
    }
}

// UPWINDALL::InterpolateInZ: wall-reduced vertical interpolation for w.
__device__ __forceinline__ double interp_z_all(const double* f, long q, long sz, double upw_lo, double upw_frac, int t) {
    if (t == ERFX_ADV_CENTERED2) return 0.5 * (f[q] + f[q - sz]);
    double sp2 = wide_stencil(t) ? f[q + 2 * sz] : 0.;
    double sp1 = f[q + sz], s = f[q], sm1 = f[q - sz], sm2 = f[q - 2 * sz];
    double sm3 = wide_stencil(t) ? f[q - 3 * sz] : 0.;
    if (upw_lo != 0.) upw_lo = (upw_lo > 0) ? 1. : -1.;
    upw_lo *= upw_frac;
    double a1 = (s + sm1), a2 = (sp1 + sm2), d1 = (s - sm1), d2 = (sp1 - sm2), a3 = (sp2 + sm3), d3 = (sp2 - sm3);
    const double g1_3_4 = (7.0 / 12.0), g2_3_4 = (1.0 / 12.0);
    const double g1_5_6 = (37.0 / 60.0), g2_5_6 = (2.0 / 15.0), g3_5_6 = (1.0 / 60.0);
    double iv = 0.0;
    if (t == ERFX_ADV_UPWIND3) iv = g1_3_4 * a1 - g2_3_4 * a2 + upw_lo * g2_3_4 * (d2 - 3.0 * d1);
    else if (t == ERFX_ADV_CENTERED4) iv = g1_3_4 * a1 - g2_3_4 * a2;
    else if (t == ERFX_ADV_UPWIND5) iv = g1_5_6 * a1 - g2_5_6 * a2 + g3_5_6 * a3 - upw_lo * g3_5_6 * (d3 - 5.0 * d2 + 10.0 * d1);
    else if (t == ERFX_ADV_CENTERED6) iv = g1_5_6 * a1 - g2_5_6 * a2 + g3_5_6 * a3;
    return iv;
}

// One box launch per precision: the functor takes (i, j, k) and a
// std::bool_constant saying whether face values are taken in mixed precision,
// so each form is its own kernel (the FP64 kernel carries no FP32 code).
template <class F>
inline void launch_box_prec(bool mixed, const Box& b, hipStream_t s, F f) {
    // the scalar and theta stencils reach several levels up and down: tiles
    // ordered k fastest (launch_box_z)
    if (mixed) launch_box_z(b, s, [=] __device__(int i, int j, int k) { f(i, j, k, std::true_type{}); });
    else launch_box_z(b, s, [=] __device__(int i, int j, int k) { f(i, j, k, std::false_type{}); });
}

// ---- momentum advection on a terrain-following grid ---------------------
// AdvectionSrcForMom with metrics (ERF_AdvectionSrcForMom.cpp:204-288 for
// C2/C2, ERF_AdvectionSrcForMom_T.H:22-330 otherwise), static terrain, unit
// map factors: the flat fluxes weighted by the face areas and the edge
// heights h_zeta, divided by detJ averaged to the face.  Each returns that
// quotient; the right-hand side is its negative.
struct TAdvData {
    const double *ru, *rv, *u, *v, *w, *om, *Ax, *Ay, *Az, *J;
    Zn z;
    long sy, sz;
    double dxi, dyi, dzi, hf, vf;
    int ht, vt, nz;
    bool c2c2;
};
template <bool mixed>      // FP32 face values (mixed precision)
struct TAdv : TAdvData {

    __device__ double x(int i, int j, int k, long q) const {
        double src;
        if (c2c2) {
            double xflux_hi = 0.25 * (ru[q] * 1.0 + ru[q + 1] * 1.0) * (u[q + 1] + u[q]) * 0.5 * (Ax[q] + Ax[q + 1]);
            double xflux_lo = 0.25 * (ru[q] * 1.0 + ru[q - 1] * 1.0) * (u[q - 1] + u[q]) * 0.5 * (Ax[q] + Ax[q - 1]);
            double met_h_zeta_yhi = met::h_zeta_EdgeK(z, i, j + 1, k, dzi);
            double yflux_hi = 0.25 * (rv[q + sy] * 1.0 + rv[q - 1 + sy] * 1.0) * (u[q + sy] + u[q]) * met_h_zeta_yhi;
            double met_h_zeta_ylo = met::h_zeta_EdgeK(z, i, j, k, dzi);
            double yflux_lo = 0.25 * (rv[q] * 1.0 + rv[q - 1] * 1.0) * (u[q - sy] + u[q]) * met_h_zeta_ylo;
            double zflux_hi = 0.25 * (om[q + sz] + om[q - 1 + sz]) * (u[q + sz] + u[q]) * 0.5 * (Az[q + sz] + Az[q - 1 + sz]);
            double zflux_lo = 0.25 * (om[q] + om[q - 1]) * (u[q - sz] + u[q]) * 0.5 * (Az[q] + Az[q - 1]);
            double mfsq = 1.0 * 1.0;
            src = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq + (zflux_hi - zflux_lo) * dzi;
        } else {
            double rho_u_avg_hi = 0.5 * (ru[q + 1] * 1.0 + ru[q] * 1.0);
            double interp_hi = interp_pick(mixed, ht, u, q + 1, 1, rho_u_avg_hi, hf);
            double rho_u_avg_lo = 0.5 * (ru[q - 1] * 1.0 + ru[q] * 1.0);
            double interp_lo = interp_pick(mixed, ht, u, q, 1, rho_u_avg_lo, hf);
            double centFluxXXNext = rho_u_avg_hi * interp_hi * 0.5 * (Ax[q] + Ax[q + 1]);
            double centFluxXXPrev = rho_u_avg_lo * interp_lo * 0.5 * (Ax[q] + Ax[q - 1]);
            double rho_v_avg_hi = 0.5 * (rv[q + sy] * 1.0 + rv[q - 1 + sy] * 1.0);
            interp_hi = interp_pick(mixed, ht, u, q + sy, sy, rho_v_avg_hi, hf);
            double rho_v_avg_lo = 0.5 * (rv[q] * 1.0 + rv[q - 1] * 1.0);
            interp_lo = interp_pick(mixed, ht, u, q, sy, rho_v_avg_lo, hf);
            double edgeFluxXYNext = rho_v_avg_hi * interp_hi * met::h_zeta_EdgeK(z, i, j + 1, k, dzi);
            double edgeFluxXYPrev = rho_v_avg_lo * interp_lo * met::h_zeta_EdgeK(z, i, j, k, dzi);
            double Omega_avg_hi = 0.5 * (om[q + sz] + om[q - 1 + sz]) * 0.5 * (Az[q + sz] + Az[q - 1 + sz]);
            double Omega_avg_lo = 0.5 * (om[q] + om[q - 1]) * 0.5 * (Az[q] + Az[q - 1]);
            interp_hi = interp_pick(mixed, vt, u, q + sz, sz, Omega_avg_hi, vf);
            interp_lo = interp_pick(mixed, vt, u, q, sz, Omega_avg_lo, vf);
            double edgeFluxXZNext = Omega_avg_hi * interp_hi;
            double edgeFluxXZPrev = Omega_avg_lo * interp_lo;
            double mfsq = 1 / (1.0 * 1.0);
            src = (centFluxXXNext - centFluxXXPrev) * dxi * mfsq + (edgeFluxXYNext - edgeFluxXYPrev) * dyi * mfsq
                + (edgeFluxXZNext - edgeFluxXZPrev) * dzi;
        }
        src /= 0.5 * (J[q] + J[q - 1]);
        return src;
    }

    __device__ double y(int i, int j, int k, long q) const {
        double src;
        if (c2c2) {
            double met_h_zeta_xhi = met::h_zeta_EdgeK(z, i + 1, j, k, dzi);
            double xflux_hi = 0.25 * (ru[q + 1] * 1.0 + ru[q + 1 - sy] * 1.0) * (v[q + 1] + v[q]) * met_h_zeta_xhi;
            double met_h_zeta_xlo = met::h_zeta_EdgeK(z, i, j, k, dzi);
            double xflux_lo = 0.25 * (ru[q] * 1.0 + ru[q - sy] * 1.0) * (v[q - 1] + v[q]) * met_h_zeta_xlo;
            double yflux_hi = 0.25 * (rv[q + sy] * 1.0 + rv[q] * 1.0) * (v[q + sy] + v[q]) * 0.5 * (Ay[q] + Ay[q + sy]);
            double yflux_lo = 0.25 * (rv[q] * 1.0 + rv[q - sy] * 1.0) * (v[q - sy] + v[q]) * 0.5 * (Ay[q] + Ay[q - sy]);
            double zflux_hi = 0.25 * (om[q + sz] + om[q - sy + sz]) * (v[q + sz] + v[q]) * 0.5 * (Az[q + sz] + Az[q - sy + sz]);
            double zflux_lo = 0.25 * (om[q] + om[q - sy]) * (v[q - sz] + v[q]) * 0.5 * (Az[q] + Az[q - sy]);
            double mfsq = 1.0 * 1.0;
            src = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq + (zflux_hi - zflux_lo) * dzi;
        } else {
            double rho_u_avg_hi = 0.5 * (ru[q + 1] * 1.0 + ru[q + 1 - sy] * 1.0);
            double rho_u_avg_lo = 0.5 * (ru[q] * 1.0 + ru[q - sy] * 1.0);
            double interp_hi = interp_pick(mixed, ht, v, q + 1, 1, rho_u_avg_hi, hf);
            double interp_lo = interp_pick(mixed, ht, v, q, 1, rho_u_avg_lo, hf);
            double edgeFluxYXNext = rho_u_avg_hi * interp_hi * met::h_zeta_EdgeK(z, i + 1, j, k, dzi);
            double edgeFluxYXPrev = rho_u_avg_lo * interp_lo * met::h_zeta_EdgeK(z, i, j, k, dzi);
            double rho_v_avg_hi = 0.5 * (rv[q] * 1.0 + rv[q + sy] * 1.0);
            double rho_v_avg_lo = 0.5 * (rv[q] * 1.0 + rv[q - sy] * 1.0);
            interp_hi = interp_pick(mixed, ht, v, q + sy, sy, rho_v_avg_hi, hf);
            interp_lo = interp_pick(mixed, ht, v, q, sy, rho_v_avg_lo, hf);
            double centFluxYYNext = rho_v_avg_hi * 0.5 * (Ay[q] + Ay[q + sy]) * interp_hi;
            double centFluxYYPrev = rho_v_avg_lo * 0.5 * (Ay[q] + Ay[q - sy]) * interp_lo;
            double Omega_avg_hi = 0.5 * (om[q + sz] + om[q - sy + sz]) * 0.5 * (Az[q + sz] + Az[q - sy + sz]);
            double Omega_avg_lo = 0.5 * (om[q] + om[q - sy]) * 0.5 * (Az[q] + Az[q - sy]);
            interp_hi = interp_pick(mixed, vt, v, q + sz, sz, Omega_avg_hi, vf);
            interp_lo = interp_pick(mixed, vt, v, q, sz, Omega_avg_lo, vf);
            double edgeFluxYZNext = Omega_avg_hi * interp_hi;
            double edgeFluxYZPrev = Omega_avg_lo * interp_lo;
            double mfsq = 1 / (1.0 * 1.0);
            src = (edgeFluxYXNext - edgeFluxYXPrev) * dxi * mfsq + (centFluxYYNext - centFluxYYPrev) * dyi * mfsq
                + (edgeFluxYZNext - edgeFluxYZPrev) * dzi;
        }
        src /= 0.5 * (J[q] + J[q - sy]);
        return src;
    }

    // interior z-faces 1 <= k <= nz-1
    __device__ double zf(int i, int j, int k, long q) const {
        const int hz = nz, lz = 0;
        double src;
        if (c2c2) {
            double met_h_zeta_xhi = met::h_zeta_EdgeJ(z, i + 1, j, k, dzi);
            double xflux_hi = 0.25 * (ru[q + 1] + ru[q + 1 - sz]) * 1.0 * (w[q + 1] + w[q]) * met_h_zeta_xhi;
            double met_h_zeta_xlo = met::h_zeta_EdgeJ(z, i, j, k, dzi);
            double xflux_lo = 0.25 * (ru[q] + ru[q - sz]) * 1.0 * (w[q - 1] + w[q]) * met_h_zeta_xlo;
            double met_h_zeta_yhi = met::h_zeta_EdgeI(z, i, j + 1, k, dzi);
            double yflux_hi = 0.25 * (rv[q + sy] + rv[q + sy - sz]) * 1.0 * (w[q + sy] + w[q]) * met_h_zeta_yhi;
            double met_h_zeta_ylo = met::h_zeta_EdgeI(z, i, j, k, dzi);
            double yflux_lo = 0.25 * (rv[q] + rv[q - sz]) * 1.0 * (w[q - sy] + w[q]) * met_h_zeta_ylo;
            double zflux_lo = 0.25 * (om[q] + om[q - sz]) * (w[q] + w[q - sz]);
            double zflux_hi = (k == hz) ? om[q] * w[q] * Az[q]
                                        : 0.25 * (om[q] + om[q + sz]) * (w[q] + w[q + sz]) * 0.5 * (Az[q] + Az[q + sz]);
            double mfsq = 1.0 * 1.0;
            src = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq + (zflux_hi - zflux_lo) * dzi;
        } else {
            double rho_u_avg_hi = 0.5 * (ru[q + 1] + ru[q + 1 - sz]) * 1.0;
            double rho_u_avg_lo = 0.5 * (ru[q] + ru[q - sz]) * 1.0;
            double interp_hi = interp_pick(mixed, ht, w, q + 1, 1, rho_u_avg_hi, hf);
            double interp_lo = interp_pick(mixed, ht, w, q, 1, rho_u_avg_lo, hf);
            double edgeFluxZXNext = rho_u_avg_hi * interp_hi * met::h_zeta_EdgeJ(z, i + 1, j, k, dzi);
            double edgeFluxZXPrev = rho_u_avg_lo * interp_lo * met::h_zeta_EdgeJ(z, i, j, k, dzi);
            double rho_v_avg_hi = 0.5 * (rv[q + sy] + rv[q + sy - sz]) * 1.0;
            interp_hi = interp_pick(mixed, ht, w, q + sy, sy, rho_v_avg_hi, hf);
            double rho_v_avg_lo = 0.5 * (rv[q] + rv[q - sz]) * 1.0;
            interp_lo = interp_pick(mixed, ht, w, q, sy, rho_v_avg_lo, hf);
            double edgeFluxZYNext = rho_v_avg_hi * interp_hi * met::h_zeta_EdgeI(z, i, j + 1, k, dzi);
            double edgeFluxZYPrev = rho_v_avg_lo * interp_lo * met::h_zeta_EdgeI(z, i, j, k, dzi);
            double Omega_avg_hi = (k == hz) ? om[q] * Az[q] : 0.25 * (om[q] + om[q + sz]) * (Az[q] + Az[q + sz]);
            double centFluxZZNext = Omega_avg_hi;
            if (k == hz) {
                centFluxZZNext *= w[q];
            } else {
                if (k == hz - 1) {
                    interp_hi = interp_z_all(w, q + sz, sz, Omega_avg_hi, 0, ERFX_ADV_CENTERED2);
                } else if (k == hz - 2 || k == lz + 1) {
                    if (vt != ERFX_ADV_CENTERED2 && vt != ERFX_ADV_UPWIND3)
                        interp_hi = interp_z_all(w, q + sz, sz, Omega_avg_hi, 0, ERFX_ADV_CENTERED4);
                    else
                        interp_hi = interp_z_all(w, q + sz, sz, Omega_avg_hi, vf, vt);
                } else {
                    interp_hi = interp_pick(mixed, vt, w, q + sz, sz, Omega_avg_hi, vf);
                }
                centFluxZZNext *= interp_hi;
            }
            double Omega_avg_lo = (k == 0) ? om[q] * Az[q] : 0.25 * (om[q] + om[q - sz]) * (Az[q] + Az[q - sz]);
            double centFluxZZPrev = Omega_avg_lo;
            if (k == 0) {
                centFluxZZPrev *= w[q];
            } else {
                if (k == lz + 1) {
                    interp_lo = interp_z_all(w, q, sz, Omega_avg_lo, 0, ERFX_ADV_CENTERED2);
                } else if (k == lz + 2 || k == hz - 1) {
                    if (vt != ERFX_ADV_CENTERED2 && vt != ERFX_ADV_UPWIND3)
                        interp_lo = interp_z_all(w, q, sz, Omega_avg_lo, 0, ERFX_ADV_CENTERED4);
                    else
                        interp_lo = interp_z_all(w, q, sz, Omega_avg_lo, vf, vt);
                } else {
                    interp_lo = interp_pick(mixed, vt, w, q, sz, Omega_avg_lo, vf);
                }
                centFluxZZPrev *= interp_lo;
            }
            double mfsq = 1.0 * 1.0;
            src = (edgeFluxZXNext - edgeFluxZXPrev) * dxi * mfsq + (edgeFluxZYNext - edgeFluxZYPrev) * dyi * mfsq
                + (centFluxZZNext - centFluxZZPrev) * dzi;
        }
        double denom = 0.5 * (J[q] + J[q - sz]);
        src /= denom;
        return src;
    }
};

// ---- strain (ComputeStrain_N, Source/Diffusion/ERF_ComputeStrain_N.cpp:28-229) ----
struct StrainBC {
    bool zl_u, zh_u, zl_v, zh_v;   // only the z walls can be Dirichlet for tangential velocity here
};

__device__ __forceinline__ double strain13(const double* u, const double* w, long q, long sz, double dxi, double dzi,
                                           bool lo_wall, bool hi_wall) {
    if (lo_wall)
        return 0.5 * ((-(8. / 3.) * u[q - sz] + 3. * u[q] - (1. / 3.) * u[q + sz]) * dzi + (w[q] - w[q - 1]) * dxi * 1.0);
    if (hi_wall)
        return 0.5 * (-(-(8. / 3.) * u[q] + 3. * u[q - sz] - (1. / 3.) * u[q - 2 * sz]) * dzi + (w[q] - w[q - 1]) * dxi * 1.0);
    return 0.5 * ((u[q] - u[q - sz]) * dzi + (w[q] - w[q - 1]) * dxi * 1.0);
}
__device__ __forceinline__ double strain23(const double* v, const double* w, long q, long sy, long sz, double dyi,
                                           double dzi, bool lo_wall, bool hi_wall) {
    if (lo_wall)
        return 0.5 * ((-(8. / 3.) * v[q - sz] + 3. * v[q] - (1. / 3.) * v[q + sz]) * dzi + (w[q] - w[q - sy]) * dyi * 1.0);
    if (hi_wall)
        return 0.5 * (-(-(8. / 3.) * v[q] + 3. * v[q - sz] - (1. / 3.) * v[q - 2 * sz]) * dzi + (w[q] - w[q - sy]) * dyi * 1.0);
    return 0.5 * ((v[q] - v[q - sz]) * dzi + (w[q] - w[q - sy]) * dyi * 1.0);
}
__device__ __forceinline__ double strain12(const double* u, const double* v, long q, long sy, double dxi, double dyi) {
    return 0.5 * ((u[q] / 1.0 - u[q - sy] / 1.0) * dyi + (v[q] / 1.0 - v[q - 1] / 1.0) * dxi) * 1.0 * 1.0;
}

static StrainBC strain_bc(const Params& P) {
    StrainBC b;
    b.zl_u = P.bck[0][2] == EXT_DIR;
    b.zh_u = P.bck[0][5] == EXT_DIR;
    b.zl_v = P.bck[1][2] == EXT_DIR;
    b.zh_v = P.bck[1][5] == EXT_DIR;
    return b;
}

static void check_lateral_strain(const Params& P) {
    // ComputeStrain_N's one-sided stencils for a Dirichlet velocity on a
    // lateral face (ERF_ComputeStrain_N.cpp:67-120) apply only where the edge
    // box starts on the domain face; make_tau_terms grows the strain boxes by
    // one laterally (ERF_make_tau_terms.cpp:135-143), so on a lateral face
    // (Inflow, no-slip) they are never taken, and the centred stencil reads
    // the ghost value there.  The z walls keep theirs (the boxes are not grown
    // in z at a wall).
    (void)P;
}

// Strain of the old velocity and ComputeTurbulentViscosityLES, Smagorinsky
// (ERF_advance_dycore.cpp:114-184; ERF_ComputeTurbulentViscosity.cpp:32-128, 230-330).
void turbulent_viscosity(const Params& P, const SlowWork& W, const double* cons, const double* u,
                         const double* v, const double* w, double* eddy, hipStream_t s) {
    check_lateral_strain(P);
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const double dxi = P.dxInv[0], dyi = P.dxInv[1], dzi = P.dxInv[2];
    const StrainBC sb = strain_bc(P);
    double *t11 = W.tau[0], *t22 = W.tau[1], *t33 = W.tau[2], *t12 = W.tau[3], *t13 = W.tau[4], *t23 = W.tau[5];
    // strain on cells grown (1,1,0) and the grown edge boxes
    launch_box(Box(-1, -1, 0, nx + 1, ny + 1, nz), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (i <= nx && j <= ny && k <= nz - 1) {
            t11[q] = (u[q + 1] / 1.0 - u[q] / 1.0) * dxi * 1.0 * 1.0;
            t22[q] = (v[q + sy] / 1.0 - v[q] / 1.0) * dyi * 1.0 * 1.0;
            t33[q] = (w[q + sz] - w[q]) * dzi;
        }
        if (k <= nz - 1) t12[q] = strain12(u, v, q, sy, dxi, dyi);
        if (j <= ny) t13[q] = strain13(u, w, q, sz, dxi, dzi, sb.zl_u && k == 0, sb.zh_u && k == nz);
        if (i <= nx) t23[q] = strain23(v, w, q, sy, sz, dyi, dzi, sb.zl_v && k == 0, sb.zh_v && k == nz);
    });
    const double Cs = P.cfg.Cs;
    const double inv_Pr_t = 1.0 / P.cfg.Pr_t;
    const double inv_Sc_t = 1.0 / P.cfg.Sc_t;
    double* mu_h = eddy;                 // EddyDiff::Mom_h
    double* mu_v = eddy + 6 * ne;        // EddyDiff::Mom_v
    launch_box(Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        double s11bar = t11[q], s22bar = t22[q], s33bar = t33[q];
        double s12bar = 0.25 * (t12[q] + t12[q + sy] + t12[q + 1] + t12[q + 1 + sy]);
        double s13bar = 0.25 * (t13[q] + t13[q + sz] + t13[q + 1] + t13[q + 1 + sz]);
        double s23bar = 0.25 * (t23[q] + t23[q + sz] + t23[q + sy] + t23[q + sy + sz]);
        double SmnSmn = s11bar * s11bar + s22bar * s22bar + s33bar * s33bar + 2.0 * s12bar * s12bar +
                        2.0 * s13bar * s13bar + 2.0 * s23bar * s23bar;
        double cellVolMsf = 1.0 / (dxi * 1.0 * dyi * 1.0 * dzi);
        double DeltaMsf = erfx_pow(cellVolMsf, 1.0 / 3.0);
        double CsDeltaSqrMsf = Cs * Cs * DeltaMsf * DeltaMsf;
        double m = CsDeltaSqrMsf * cons[q] * sqrt(2.0 * SmnSmn);
        mu_h[q] = m;
        mu_v[q] = m;
    });
    // lateral extrapolation of Mom_h/Mom_v into one ghost layer, x then y planes
    launch_box(Box(0, -1, 0, 1, ny, nz - 1), s, [=] __device__(int side, int j, int k) {
        int ig = side == 0 ? -1 : nx, is = side == 0 ? 0 : nx - 1;
        int lj = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
        long d = L.ix(ig, j, k), sidx = L.ix(is, lj, k);
        mu_h[d] = mu_h[sidx];
        mu_v[d] = mu_v[sidx];
    });
    launch_box(Box(-1, 0, 0, nx, 1, nz - 1), s, [=] __device__(int i, int side, int k) {
        int jg = side == 0 ? -1 : ny, js = side == 0 ? 0 : ny - 1;
        int li = i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i);
        long d = L.ix(i, jg, k), sidx = L.ix(li, js, k);
        mu_h[d] = mu_h[sidx];
        mu_v[d] = mu_v[sidx];
    });
    // Theta, Scalar and Q diffusivities (comps 1, 4, 5 and their _v)
    launch_box(Box(-1, -1, 0, nx, ny, nz - 1), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        double m = mu_h[q];
        double th = m * inv_Pr_t;
        eddy[q + 1 * ne] = th; eddy[q + 7 * ne] = th;
        double sc = m * inv_Sc_t;
        eddy[q + 4 * ne] = sc; eddy[q + 10 * ne] = sc;
        eddy[q + 5 * ne] = sc; eddy[q + 11 * ne] = sc;
    });
    fill_boundary_periodic(P, eddy, 13, 1, -1, s);
    // vertical extrapolation into one ghost layer for comps 0,1,4,5 (+6)
    launch_box(Box(-1, -1, 0, nx, ny, 1), s, [=] __device__(int i, int j, int side) {
        int kg = side == 0 ? -1 : nz, ks = side == 0 ? 0 : nz - 1;
        long d = L.ix(i, j, kg), sidx = L.ix(i, j, ks);
        const int comps[8] = {0, 6, 1, 7, 4, 10, 5, 11};
        for (int c = 0; c < 8; ++c) eddy[d + comps[c] * ne] = eddy[sidx + comps[c] * ne];
    });
}

// erf_make_tau_terms, flat (Source/TimeIntegration/ERF_make_tau_terms.cpp:216-303):
// expansion rate + strain of the stage velocity, then ComputeStress{Cons,Var}Visc_N
// (Source/Diffusion/ERF_ComputeStress_N.cpp:1-189), written straight into tau.
static void make_tau_terms(const Params& P, const SlowWork& W, const double* cons, const double* u,
                           const double* v, const double* w, const double* eddy, hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const double dxi = P.dxInv[0], dyi = P.dxInv[1], dzi = P.dxInv[2];
    const StrainBC sb = strain_bc(P);
    const bool consA = P.consA, turb = P.use_turb;
    const double mu_eff = consA ? 2.0 * P.cfg.dynamic_viscosity / P.cfg.rho0_trans : 2.0 * P.cfg.dynamic_viscosity;
    const double OneThird = (1. / 3.);
    double *t11 = W.tau[0], *t22 = W.tau[1], *t33 = W.tau[2], *t12 = W.tau[3], *t13 = W.tau[4], *t23 = W.tau[5];
    const double* mh = eddy;
    const double* mv = eddy ? eddy + 6 * ne : nullptr;
    // cells grown (1,1,0) and edges of the valid nodal boxes in one launch
    launch_box(Box(-1, -1, 0, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (k <= nz - 1) {
            double er = (u[q + 1] / 1.0 - u[q] / 1.0) * dxi * (1.0 * 1.0) +
                        (v[q + sy] / 1.0 - v[q] / 1.0) * dyi * (1.0 * 1.0) + (w[q + sz] - w[q]) * dzi;
            double s11 = (u[q + 1] / 1.0 - u[q] / 1.0) * dxi * 1.0 * 1.0;
            double s22 = (v[q + sy] / 1.0 - v[q] / 1.0) * dyi * 1.0 * 1.0;
            double s33 = (w[q + sz] - w[q]) * dzi;
            double m11, m33;
            if (!turb) {
                m11 = consA ? cons[q] * mu_eff : mu_eff;
                m33 = m11;
            } else {
                double base = consA ? cons[q] * mu_eff : mu_eff;
                m11 = base + 2.0 * mh[q];
                m33 = base + 2.0 * mv[q];
            }
            t11[q] = -m11 * (s11 - OneThird * er);
            t22[q] = -m11 * (s22 - OneThird * er);
            t33[q] = -m33 * (s33 - OneThird * er);
        }
        if (i < 0 || j < 0) return;
        // edges on nodal valid boxes
        if (k <= nz - 1) {                                   // xy edges i in [0,nx], j in [0,ny]
            double s12 = strain12(u, v, q, sy, dxi, dyi);
            double mu;
            double rho_bar = consA ? 0.25 * (cons[q - 1] + cons[q] + cons[q - 1 - sy] + cons[q - sy]) : 0.0;
            if (turb) {
                double mu_bar = 0.25 * (mh[q - 1] + mh[q] + mh[q - 1 - sy] + mh[q - sy]);
                mu = (consA ? rho_bar * mu_eff : mu_eff) + 2.0 * mu_bar;
                t12[q] = s12 * -mu;
            } else {
                t12[q] = consA ? s12 * (-rho_bar * mu_eff) : s12 * -mu_eff;
            }
        }
        if (j <= ny - 1) {                                   // xz edges i in [0,nx], k in [0,nz]
            double s13 = strain13(u, w, q, sz, dxi, dzi, sb.zl_u && k == 0, sb.zh_u && k == nz);
            double rho_bar = consA ? 0.25 * (cons[q - 1] + cons[q] + cons[q - 1 - sz] + cons[q - sz]) : 0.0;
            if (turb) {
                double mu_bar = 0.25 * (mv[q - 1] + mv[q] + mv[q - 1 - sz] + mv[q - sz]);
                double mu = (consA ? rho_bar * mu_eff : mu_eff) + 2.0 * mu_bar;
                t13[q] = s13 * -mu;
            } else {
                t13[q] = consA ? s13 * (-rho_bar * mu_eff) : s13 * -mu_eff;
            }
        }
        if (i <= nx - 1) {                                   // yz edges j in [0,ny], k in [0,nz]
            double s23 = strain23(v, w, q, sy, sz, dyi, dzi, sb.zl_v && k == 0, sb.zh_v && k == nz);
            double rho_bar = consA ? 0.25 * (cons[q - sy] + cons[q] + cons[q - sy - sz] + cons[q - sz]) : 0.0;
            if (turb) {
                double mu_bar = 0.25 * (mv[q - sy] + mv[q] + mv[q - sy - sz] + mv[q - sz]);
                double mu = (consA ? rho_bar * mu_eff : mu_eff) + 2.0 * mu_bar;
                t23[q] = s23 * -mu;
            } else {
                t23[q] = consA ? s23 * (-rho_bar * mu_eff) : s23 * -mu_eff;
            }
        }
    });
}

// Diffusive flux of the scalar `prim_index` through the lo face of q along
// stride st (DiffusionSrcForState_N, ERF_DiffusionSrcForState_N.cpp:197-600).
__device__ __forceinline__ double diff_flux(const double* cons, const double* cp, const double* mu, long q, long st,
                                            double dinv, double ae, bool consA, bool turb) {
    double rhoAlpha;
    if (consA) {
        double rhoFace = 0.5 * (cons[q] + cons[q - st]);
        rhoAlpha = rhoFace * ae;
    } else {
        rhoAlpha = ae;
    }
    if (turb) rhoAlpha += 0.5 * (mu[q] + mu[q - st]);
    return -rhoAlpha * (cp[q] - cp[q - st]) * dinv;
}

// The same flux on a face where the scalar is Dirichlet (an Inflow face): the
// one-sided stencils through the face value held in the ghost cell
// (ERF_DiffusionSrcForState_N.cpp:208-227), side -1 on the domain's lo face,
// +1 on its hi face, 0 elsewhere.
__device__ __forceinline__ double diff_flux_bc(const double* cons, const double* cp, const double* mu, long q,
                                               long st, double dinv, double ae, bool consA, bool turb, int side) {
    if (side == 0) return diff_flux(cons, cp, mu, q, st, dinv, ae, consA, turb);
    double rhoAlpha;
    if (consA) {
        double rhoFace = 0.5 * (cons[q] + cons[q - st]);
        rhoAlpha = rhoFace * ae;
    } else {
        rhoAlpha = ae;
    }
    if (turb) rhoAlpha += 0.5 * (mu[q] + mu[q - st]);
    if (side < 0) return -rhoAlpha * (-(8. / 3.) * cp[q - st] + 3. * cp[q] - (1. / 3.) * cp[q + st]) * dinv;
    return -rhoAlpha * ((8. / 3.) * cp[q] - 3. * cp[q - st] + (1. / 3.) * cp[q - 2 * st]) * dinv;
}

// Departure of cell q (level k) from the level's plane averages in the moist
// buoyancy of types 2, 3 and 4 (ERF_make_buoyancy.cpp:272-323, 347-372).
// pavg rows of nz values: rho, theta, qv, qc, qp averages, then the
// temperature of the averages (PLANE_T).
__device__ __forceinline__ double moist_buoyancy_dev(int btype, int nq, const double* cons, const double* pq1,
                                                     const double* pq2, const double* pq3, const double* pavg,
                                                     int nz, long ne, long q, int k) {
    const double rho_d = pavg[k], theta_d = pavg[nz + k], qv_d = pavg[2 * nz + k];
    const double qc_d = pavg[3 * nz + k], qp_d = pavg[4 * nz + k];
    const double temp1d = pavg[PLANE_T * nz + k];
    const double rho = cons[q], rt = cons[q + ne];
    const double temp3d = eos_T(rho, rt, cons[q + ERFX_RHOQ1 * ne] / rho);
    const double qv = (nq >= 1) ? pq1[q] : 0.0;
    const double qc = (nq >= 2) ? pq2[q] : 0.0;
    const double qp = (nq >= 3) ? pq3[q] : 0.0;
    (void)rho_d;
    if (btype == 2)
        return 0.61 * (qv - qv_d) - (qc - qc_d + qp - qp_d) +
               (temp3d - temp1d) / temp1d * (1.0 + 0.61 * qv_d - qc_d - qp_d);
    if (btype == 4)
        return 0.61 * (qv - qv_d) - (qc - qc_d + qp - qp_d) + (rt / rho - theta_d) / theta_d;
    return 0.61 * qv - (qc + qp) + (temp3d - temp1d) / temp1d;
}

// erf_slow_rhs_pre with make_mom_sources (ERF_TI_slow_rhs_fun.H:216-241), flat.
void slow_rhs_pre(const Params& P, const SlowWork& W, int nrk, double slow_dt, const erfx_state& rhs,
                  const erfx_state& data, const double* prim, const erfx_state& scratch, const double* u,
                  const double* v, const double* w, const double* eddy, const double* r0, const double* p0,
                  hipStream_t s) {
    (void)nrk; (void)slow_dt;
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const double dxi = P.dxInv[0], dyi = P.dxInv[1], dzi = P.dxInv[2];
    const bool use_diff = P.use_diff, consA = P.consA, turb = P.use_turb;
    if (use_diff) {
        if (P.terrain) {   // make_tau_terms' terrain branch; DiffusionSrcForState_T of theta
            make_tau_terms_T(P, W, data.cons, u, v, w, s);
            diffusion_state_T(P, W, 1, 1, data.cons, prim, s);
        } else {
            make_tau_terms(P, W, data.cons, u, v, w, eddy, s);
        }
    }

    // p' on cells grown by one, k >= 0 (ERF_slow_rhs_pre.cpp:337-358)
    double* pp = W.pp;
    const double* rt = data.cons + ne;
    const bool moist = P.moist;
    if (!moist) {
        launch_box(Box(-1, -1, 0, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
            const long q = L.ix(i, j, k);
            pp[q] = eos_p_of_rt(rt[q]) - p0[q];
        });
    } else {
        const double* rq1 = data.cons + ERFX_RHOQ1 * ne;
        const double* rho = data.cons;
        launch_box(Box(-1, -1, 0, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
            const long q = L.ix(i, j, k);
            double qv_for_p = rq1[q] / rho[q];
            pp[q] = eos_p_of_rt_qv(rt[q], qv_for_p) - p0[q];
        });
        // plane averages and the level temperatures of the moist buoyancy (types 2-4)
        if (P.cfg.buoyancy_type != 1) plane_averages(P, data.cons, prim, W.pavg_part, W.pavg, s);
    }

    // cell RHS of rho and rho*theta: AdvectionSrcForRho, AdvectionSrcForScalars
    // (theta, dycore scheme) and DiffusionSrcForState_N (theta)
    const int ht = P.cfg.dycore_horiz_adv, vt = P.cfg.dycore_vert_adv;
    const double hf = P.cfg.dycore_horiz_upw_frac, vf = P.cfg.dycore_vert_upw_frac;
    const bool c2c2 = (ht == ERFX_ADV_CENTERED2 && vt == ERFX_ADV_CENTERED2);
    const bool mixed_rt = P.cfg.precision == ERFX_PREC_MIXED;
    double ae_theta = consA ? P.cfg.alpha_T : P.cfg.rho0_trans * P.cfg.alpha_T;
    // Dirichlet theta faces (Inflow): one-sided diffusive fluxes there
    const bool dxlo = P.bck[3][0] == EXT_DIR, dxhi = P.bck[3][3] == EXT_DIR;
    const bool dylo = P.bck[3][1] == EXT_DIR, dyhi = P.bck[3][4] == EXT_DIR;
    const double* mu_th_h = eddy ? eddy + 1 * ne : nullptr;
    const double* mu_th_v = eddy ? eddy + 7 * ne : nullptr;
    const bool sflux = P.cfg.use_surface_heat_flux != 0;
    const double sflux_val = P.cfg.surface_heat_flux;
    double* cr0 = rhs.cons;
    double* cr1 = rhs.cons + ne;
    const double* ray = P.ray;
    const int n1 = nz + 1;
    const bool rayT = P.cfg.rayleigh_damp_T && ray;
    const bool rayU = P.cfg.rayleigh_damp_U && ray, rayV = P.cfg.rayleigh_damp_V && ray;
    const bool rayW = P.cfg.rayleigh_damp_W && ray;
    const RayAt RA{ray, P.terrain ? P.ray_pt : nullptr, n1, ne};
    const double* ru = data.xmom;
    const double* rv = data.ymom;
    const double* rw = data.zmom;
    if (P.terrain) {
        // Omega on z-faces grown by one (ERF_slow_rhs_pre.cpp:367-414): 0 at and
        // below the bottom face, rho*w above the top face, OmegaFromW between
        // (the top face included); the area-weighted fluxes of
        // AdvectionSrcForRho (ERF_AdvectionSrcForState.cpp:29-93) seed the
        // time-averaged momenta
        const Zn z{P.z_nd, L};
        const Fld fru{ru, L}, frv{rv, L};
        const double* Ax = P.ax; const double* Ay = P.ay; const double* Az = P.az; const double* Jc = P.detJ;
        const bool tdiff = use_diff;
        const double* dst_theta = use_diff ? W.dstate + ne : nullptr;
        double* omw = W.omega;
        double* sxm = scratch.xmom; double* sym = scratch.ymom; double* szm = scratch.zmom;
        launch_box(Box(-1, -1, -1, nx, ny, nz + 1), s, [=] __device__(int i, int j, int k) {
            const long q = L.ix(i, j, k);
            double o;
            if (k <= 0) o = 0.;
            else if (k == nz + 1) o = rw[q];
            else o = met::omega_from_w(i, j, k, rw[q], fru, frv, z, dxi, dyi);
            omw[q] = o;
            if (i >= 0 && j >= 0 && j <= ny - 1 && k >= 0 && k <= nz - 1) sxm[q] = Ax[q] * ru[q] / 1.0;
            if (i >= 0 && i <= nx - 1 && j >= 0 && k >= 0 && k <= nz - 1) sym[q] = Ay[q] * rv[q] / 1.0;
            if (i >= 0 && i <= nx - 1 && j >= 0 && j <= ny - 1 && k >= 0 && k <= nz) szm[q] = Az[q] * o / (1.0 * 1.0);
        });
        // cell RHS of rho (AdvectionSrcForRho) and rho*theta (AdvectionSrcForScalars
        // with 1/detJ, :298), from the area-weighted fluxes
        launch_box_prec(mixed_rt, Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k, auto mc) {
            constexpr bool mixed = decltype(mc)::value;
            const long q = L.ix(i, j, k);
            double fx0 = sxm[q], fx1 = sxm[q + 1], fy0 = sym[q], fy1 = sym[q + sy], fz0 = szm[q], fz1 = szm[q + sz];
            const double mfsq = 1.0 * 1.0;
            const double dJ = Jc[q];
            if (dJ > 0.)
                cr0[q] = -mfsq / dJ * ((fx1 - fx0) * dxi + (fy1 - fy0) * dyi + (fz1 - fz0) * dzi);
            else
                cr0[q] = 0.;
            double gx0, gx1, gy0, gy1, gz0, gz1;
            if (c2c2) {
                gx0 = fx0 * (0.5 * (prim[q] + prim[q - 1]));
                gx1 = fx1 * (0.5 * (prim[q + 1] + prim[q]));
                gy0 = fy0 * (0.5 * (prim[q] + prim[q - sy]));
                gy1 = fy1 * (0.5 * (prim[q + sy] + prim[q]));
                gz0 = fz0 * (0.5 * (prim[q] + prim[q - sz]));
                gz1 = fz1 * (0.5 * (prim[q + sz] + prim[q]));
            } else {
                gx0 = fx0 * interp_pick(mixed, ht, prim, q, 1, fx0, hf);
                gx1 = fx1 * interp_pick(mixed, ht, prim, q + 1, 1, fx1, hf);
                gy0 = fy0 * interp_pick(mixed, ht, prim, q, sy, fy0, hf);
                gy1 = fy1 * interp_pick(mixed, ht, prim, q + sy, sy, fy1, hf);
                gz0 = fz0 * interp_pick(mixed, vt, prim, q, sz, fz0, vf);
                gz1 = fz1 * interp_pick(mixed, vt, prim, q + sz, sz, fz1, vf);
            }
            double invdetJ = (dJ > 0.) ? 1. / dJ : 1.;
            double r1 = -invdetJ * mfsq * ((gx1 - gx0) * dxi + (gy1 - gy0) * dyi + (gz1 - gz0) * dzi);
            if (tdiff) r1 -= dst_theta[q];   // DiffusionSrcForState_T (made above)
            double src = 0.0;
            if (rayT && RA.act_cc(q, k)) {   // ERF_make_sources.cpp:203-225 with z_cc
                double theta = prim[q];
                src -= RA.fac_cc(q, k) * (theta - ray[RAY_THETABAR * n1 + k]) * data.cons[q];
            }
            r1 += src;
            cr1[q] = r1;
        });
    } else {
        launch_box_prec(mixed_rt, Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k, auto mc) {
            constexpr bool mixed = decltype(mc)::value;
            const long q = L.ix(i, j, k);
            double fx0 = ru[q], fx1 = ru[q + 1], fy0 = rv[q], fy1 = rv[q + sy], fz0 = rw[q], fz1 = rw[q + sz];
            cr0[q] = -(1.0 * 1.0) / 1.0 * ((fx1 - fx0) * dxi + (fy1 - fy0) * dyi + (fz1 - fz0) * dzi);
            double gx0, gx1, gy0, gy1, gz0, gz1;
            if (c2c2) {
                gx0 = fx0 * (0.5 * (prim[q] + prim[q - 1]));
                gx1 = fx1 * (0.5 * (prim[q + 1] + prim[q]));
                gy0 = fy0 * (0.5 * (prim[q] + prim[q - sy]));
                gy1 = fy1 * (0.5 * (prim[q + sy] + prim[q]));
                gz0 = fz0 * (0.5 * (prim[q] + prim[q - sz]));
                gz1 = fz1 * (0.5 * (prim[q + sz] + prim[q]));
            } else {
                gx0 = fx0 * interp_pick(mixed, ht, prim, q, 1, fx0, hf);
                gx1 = fx1 * interp_pick(mixed, ht, prim, q + 1, 1, fx1, hf);
                gy0 = fy0 * interp_pick(mixed, ht, prim, q, sy, fy0, hf);
                gy1 = fy1 * interp_pick(mixed, ht, prim, q + sy, sy, fy1, hf);
                gz0 = fz0 * interp_pick(mixed, vt, prim, q, sz, fz0, vf);
                gz1 = fz1 * interp_pick(mixed, vt, prim, q + sz, sz, fz1, vf);
            }
            double r1 = -1. * (1.0 * 1.0) * ((gx1 - gx0) * dxi + (gy1 - gy0) * dyi + (gz1 - gz0) * dzi);
            if (use_diff) {
                const double* c = data.cons;
                const int dsx0 = (dxlo && i == 0) ? -1 : 0, dsx1 = (dxhi && i == nx - 1) ? 1 : 0;
                const int dsy0 = (dylo && j == 0) ? -1 : 0, dsy1 = (dyhi && j == ny - 1) ? 1 : 0;
                double dfx0 = diff_flux_bc(c, prim, mu_th_h, q, 1, dxi, ae_theta, consA, turb, dsx0) * 1.0;
                double dfx1 = diff_flux_bc(c, prim, mu_th_h, q + 1, 1, dxi, ae_theta, consA, turb, dsx1) * 1.0;
                double dfy0 = diff_flux_bc(c, prim, mu_th_h, q, sy, dyi, ae_theta, consA, turb, dsy0) * 1.0;
                double dfy1 = diff_flux_bc(c, prim, mu_th_h, q + sy, sy, dyi, ae_theta, consA, turb, dsy1) * 1.0;
                double dfz0 = (sflux && k == 0) ? sflux_val : diff_flux(c, prim, mu_th_v, q, sz, dzi, ae_theta, consA, turb);
                double dfz1 = diff_flux(c, prim, mu_th_v, q + sz, sz, dzi, ae_theta, consA, turb);
                r1 -= (dfx1 - dfx0) * dxi * 1.0 + (dfy1 - dfy0) * dyi * 1.0 + (dfz1 - dfz0) * dzi;
            }
            // cell source of rho*theta (make_sources: Rayleigh damping, ERF_make_sources.cpp:203-225),
            // added the way ERF adds cc_src (ERF_slow_rhs_pre.cpp:521-526)
            double src = 0.0;
            if (rayT && RA.act_cc(q, k)) {
                double theta = prim[q];
                src -= RA.fac_cc(q, k) * (theta - ray[RAY_THETABAR * n1 + k]) * data.cons[q];
            }
            r1 += src;
            cr1[q] = r1;
        });

        // time-averaged momenta seeded with the stage momenta (ERF_AdvectionSrcForState.cpp:56-71)
        launch_box(Box(0, 0, 0, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
            const long q = L.ix(i, j, k);
            if (j <= ny - 1 && k <= nz - 1) scratch.xmom[q] = 1.0 * ru[q] / 1.0;
            if (i <= nx - 1 && k <= nz - 1) scratch.ymom[q] = 1.0 * rv[q] / 1.0;
            if (i <= nx - 1 && j <= ny - 1) scratch.zmom[q] = 1.0 * rw[q] / (1.0 * 1.0);
        });

    }

    // momentum RHS (AdvectionSrcForMom flat, DiffusionSrcForMom_N, pressure
    // gradient, make_mom_sources), one launch over the three face boxes
    const double* cons = data.cons;
    const double* t11 = W.tau[0]; const double* t22 = W.tau[1]; const double* t33 = W.tau[2];
    const double* t12 = W.tau[3]; const double* t13 = W.tau[4]; const double* t23 = W.tau[5];
    const double* t21 = W.tau_t[0]; const double* t31 = W.tau_t[1]; const double* t32 = W.tau_t[2];
    const double* Jd = P.detJ;
    const bool cor = P.cfg.use_coriolis != 0;
    const double cf = P.cfg.coriolis_factor, sinphi = P.cfg.sinphi, cosphi = P.cfg.cosphi;
    const double g0 = P.cfg.abl_geo_forcing[0], g1 = P.cfg.abl_geo_forcing[1], g2 = P.cfg.abl_geo_forcing[2];
    const double grav_z = -P.gravity;
    const bool zero_xlo = P.bck[0][0] == EXT_DIR, zero_xhi = P.bck[0][3] == EXT_DIR;
    const bool zero_ylo = P.bck[1][1] == EXT_DIR, zero_yhi = P.bck[1][4] == EXT_DIR;
    double* rur = rhs.xmom;
    double* rvr = rhs.ymom;
    double* rwr = rhs.zmom;
    const double* om = P.terrain ? W.omega : rw;   // Omega = rho*w on a flat grid
    const bool terr = P.terrain;
    const TAdvData TD{ru, rv, u, v, w, om, P.ax, P.ay, P.az, P.detJ, Zn{P.z_nd, L}, sy, sz, dxi, dyi, dzi, hf, vf,
                      ht, vt, nz, c2c2};
    const Zn zt{P.z_nd, L};
    const double* mrx = P.mrx;
    const double* mry = P.mry;
    const double* pq1 = prim + PRIM_Q1 * ne;
    const double* pq2 = prim + PRIM_Q2 * ne;
    const double* pq3 = prim + PRIM_Q3 * ne;
    const int btype = moist ? P.cfg.buoyancy_type : 1;
    const int nqs = P.n_qstate;
    const double* pavg = W.pavg;
    // the PressureGradient driver (ERF_slow_rhs_pre.cpp:592, 630, 679, 739)
    const double apg0 = P.cfg.abl_pressure_grad[0], apg1 = P.cfg.abl_pressure_grad[1];
    const double apg2 = P.cfg.abl_pressure_grad[2];
    // momentum face values stay FP64 in both forms: this kernel streams at
    // near the HBM peak, and FP32 face values only add registers to it.
    // Its stencils reach three levels up and down: tiles ordered k fastest.
    // The advection type is a compile-time constant when both directions use
    // U3 or U5 (tc; -1: the types at run time), so the face stencils compile
    // to straight-line code.
    const int ht_rt = ht, vt_rt = vt;
    auto body = [=] __device__(int i, int j, int k, auto tc) {
        constexpr int TT = decltype(tc)::value;
        const int ht = TT >= 0 ? TT : ht_rt;
        const int vt = TT >= 0 ? TT : vt_rt;
        constexpr bool mixed = false;
        TAdvData td = TD;
        td.ht = ht;
        td.vt = vt;
        const TAdv<mixed> TA{td};
        const long q = L.ix(i, j, k);
        // ---- x-momentum on x-faces ----
        if (j <= ny - 1 && k <= nz - 1) {
            double src;
            if (terr) {
                src = TA.x(i, j, k, q);
            } else if (c2c2) {
                double xflux_hi = 0.25 * (ru[q] * 1.0 + ru[q + 1] * 1.0) * (u[q + 1] + u[q]);
                double xflux_lo = 0.25 * (ru[q] * 1.0 + ru[q - 1] * 1.0) * (u[q - 1] + u[q]);
                double yflux_hi = 0.25 * (rv[q + sy] * 1.0 + rv[q - 1 + sy] * 1.0) * (u[q + sy] + u[q]);
                double yflux_lo = 0.25 * (rv[q] * 1.0 + rv[q - 1] * 1.0) * (u[q - sy] + u[q]);
                double zflux_hi = 0.25 * (om[q + sz] + om[q - 1 + sz]) * (u[q + sz] + u[q]);
                double zflux_lo = 0.25 * (om[q] + om[q - 1]) * (u[q - sz] + u[q]);
                src = (xflux_hi - xflux_lo) * dxi * (1.0 * 1.0) + (yflux_hi - yflux_lo) * dyi * (1.0 * 1.0) +
                      (zflux_hi - zflux_lo) * dzi;
            } else {
                double a_hi = 0.5 * (ru[q + 1] * 1.0 + ru[q] * 1.0);
                double xflux_hi = a_hi * interp_pick(mixed, ht, u, q + 1, 1, a_hi, hf);
                double a_lo = 0.5 * (ru[q - 1] * 1.0 + ru[q] * 1.0);
                double xflux_lo = a_lo * interp_pick(mixed, ht, u, q, 1, a_lo, hf);
                double b_hi = 0.5 * (rv[q + sy] * 1.0 + rv[q - 1 + sy] * 1.0);
                double yflux_hi = b_hi * interp_pick(mixed, ht, u, q + sy, sy, b_hi, hf);
                double b_lo = 0.5 * (rv[q] * 1.0 + rv[q - 1] * 1.0);
                double yflux_lo = b_lo * interp_pick(mixed, ht, u, q, sy, b_lo, hf);
                double c_hi = 0.5 * (om[q + sz] + om[q - 1 + sz]);
                double zflux_hi = c_hi * interp_pick(mixed, vt, u, q + sz, sz, c_hi, vf);
                double c_lo = 0.5 * (om[q] + om[q - 1]);
                double zflux_lo = c_lo * interp_pick(mixed, vt, u, q, sz, c_lo, vf);
                double mfsq = 1 / (1.0 * 1.0);
                src = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq +
                      (zflux_hi - zflux_lo) * dzi;
            }
            double r = -src;
            if (use_diff && terr) {   // DiffusionSrcForMom_T (ERF_DiffusionSrcForMom_T.cpp:50-58)
                double diffContrib = ((t11[q] - t11[q - 1]) * dxi * 1.0 + (t12[q + sy] - t12[q]) * dyi * 1.0 +
                                      (t13[q + sz] - t13[q]) * dzi);
                diffContrib /= 0.5 * (Jd[q] + Jd[q - 1]);
                r -= diffContrib;
            } else if (use_diff)
                r -= ((t11[q] - t11[q - 1]) * dxi * 1.0 + (t12[q + sy] - t12[q]) * dyi * 1.0 +
                      (t13[q + sz] - t13[q]) * dzi);
            double xs = 0.0;
            if (cor) {
                double rho_v_loc = 0.25 * (rv[q + sy] + rv[q] + rv[q - 1 + sy] + rv[q - 1]);
                double rho_w_loc = 0.25 * (rw[q + sz] + rw[q] + rw[q - sy + sz] + rw[q - sy]);
                xs += cf * (rho_v_loc * sinphi - rho_w_loc * cosphi);
            }
            if (rayU && RA.act_cc(q, k)) {
                double rho_on_u_face = 0.5 * (cons[q] + cons[q - 1]);
                double uu = ru[q] / rho_on_u_face;
                xs -= RA.fac_cc(q, k) * (uu - ray[RAY_UBAR * n1 + k]) * rho_on_u_face;
            }
            xs += 0.5 * (cons[q] + cons[q - 1]) * g0;
            double gpx = dxi * (pp[q] - pp[q - 1]);
            if (terr) {   // ERF_slow_rhs_pre.cpp:600-620
                double gp_zeta_on_iface;
                if (k == 0)
                    gp_zeta_on_iface = 0.5 * dzi * (pp[q - 1 + sz] + pp[q + sz] - pp[q - 1] - pp[q]);
                else if (k == nz - 1)
                    gp_zeta_on_iface = 0.5 * dzi * (pp[q - 1] + pp[q] - pp[q - 1 - sz] - pp[q - sz]);
                else
                    gp_zeta_on_iface = 0.25 * dzi * (pp[q - 1 + sz] + pp[q + sz] - pp[q - 1 - sz] - pp[q - sz]);
                gpx -= mrx[q] * gp_zeta_on_iface;   // (h_xi / h_zeta) made once
            }
            gpx *= 1.0;
            double qf = moist ? face_q(pq1, pq2, q, q - 1) : 0.0;
            r += (-gpx - apg0) / (1.0 + qf) + xs;
            if ((zero_xlo && i == 0) || (zero_xhi && i == nx)) r = 0.;
            rur[q] = r;
        }
        // ---- y-momentum on y-faces ----
        if (i <= nx - 1 && k <= nz - 1) {
            double src;
            if (terr) {
                src = TA.y(i, j, k, q);
            } else if (c2c2) {
                double xflux_hi = 0.25 * (ru[q + 1] * 1.0 + ru[q + 1 - sy] * 1.0) * (v[q + 1] + v[q]);
                double xflux_lo = 0.25 * (ru[q] * 1.0 + ru[q - sy] * 1.0) * (v[q - 1] + v[q]);
                double yflux_hi = 0.25 * (rv[q + sy] * 1.0 + rv[q] * 1.0) * (v[q + sy] + v[q]);
                double yflux_lo = 0.25 * (rv[q] * 1.0 + rv[q - sy] * 1.0) * (v[q - sy] + v[q]);
                double zflux_hi = 0.25 * (om[q + sz] + om[q - sy + sz]) * (v[q + sz] + v[q]);
                double zflux_lo = 0.25 * (om[q] + om[q - sy]) * (v[q - sz] + v[q]);
                src = (xflux_hi - xflux_lo) * dxi * (1.0 * 1.0) + (yflux_hi - yflux_lo) * dyi * (1.0 * 1.0) +
                      (zflux_hi - zflux_lo) * dzi;
            } else {
                double a_hi = 0.5 * (ru[q + 1] * 1.0 + ru[q + 1 - sy] * 1.0);
                double xflux_hi = a_hi * interp_pick(mixed, ht, v, q + 1, 1, a_hi, hf);
                double a_lo = 0.5 * (ru[q] * 1.0 + ru[q - sy] * 1.0);
                double xflux_lo = a_lo * interp_pick(mixed, ht, v, q, 1, a_lo, hf);
                double b_hi = 0.5 * (rv[q] * 1.0 + rv[q + sy] * 1.0);
                double yflux_hi = b_hi * interp_pick(mixed, ht, v, q + sy, sy, b_hi, hf);
                double b_lo = 0.5 * (rv[q] * 1.0 + rv[q - sy] * 1.0);
                double yflux_lo = b_lo * interp_pick(mixed, ht, v, q, sy, b_lo, hf);
                double c_hi = 0.5 * (om[q + sz] + om[q - sy + sz]);
                double zflux_hi = c_hi * interp_pick(mixed, vt, v, q + sz, sz, c_hi, vf);
                double c_lo = 0.5 * (om[q] + om[q - sy]);
                double zflux_lo = c_lo * interp_pick(mixed, vt, v, q, sz, c_lo, vf);
                double mfsq = 1 / (1.0 * 1.0);
                src = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq +
                      (zflux_hi - zflux_lo) * dzi;
            }
            double r = -src;
            if (use_diff && terr) {   // :60-68, tau21 in place of tau12
                double diffContrib = ((t21[q + 1] - t21[q]) * dxi * 1.0 + (t22[q] - t22[q - sy]) * dyi * 1.0 +
                                      (t23[q + sz] - t23[q]) * dzi);
                diffContrib /= 0.5 * (Jd[q] + Jd[q - sy]);
                r -= diffContrib;
            } else if (use_diff)
                r -= ((t12[q + 1] - t12[q]) * dxi * 1.0 + (t22[q] - t22[q - sy]) * dyi * 1.0 +
                      (t23[q + sz] - t23[q]) * dzi);
            double ys = 0.0;
            if (cor) {
                double rho_u_loc = 0.25 * (ru[q + 1] + ru[q] + ru[q + 1 - sy] + ru[q - sy]);
                ys += -cf * rho_u_loc * sinphi;
            }
            if (rayV && RA.act_cc(q, k)) {
                double rho_on_v_face = 0.5 * (cons[q] + cons[q - sy]);
                double vv = rv[q] / rho_on_v_face;
                ys -= RA.fac_cc(q, k) * (vv - ray[RAY_VBAR * n1 + k]) * rho_on_v_face;
            }
            ys += 0.5 * (cons[q] + cons[q - sy]) * g1;
            double gpy = dyi * (pp[q] - pp[q - sy]);
            if (terr) {   // ERF_slow_rhs_pre.cpp:650-669
                double gp_zeta_on_jface;
                if (k == 0)
                    gp_zeta_on_jface = 0.5 * dzi * (pp[q + sz] + pp[q - sy + sz] - pp[q] - pp[q - sy]);
                else if (k == nz - 1)
                    gp_zeta_on_jface = 0.5 * dzi * (pp[q] + pp[q - sy] - pp[q - sz] - pp[q - sy - sz]);
                else
                    gp_zeta_on_jface = 0.25 * dzi * (pp[q + sz] + pp[q - sy + sz] - pp[q - sz] - pp[q - sy - sz]);
                gpy -= mry[q] * gp_zeta_on_jface;   // (h_eta / h_zeta) made once
            }
            gpy *= 1.0;
            double qf = moist ? face_q(pq1, pq2, q, q - sy) : 0.0;
            r += (-gpy - apg1) / (1.0 + qf) + ys;
            if ((zero_ylo && j == 0) || (zero_yhi && j == ny)) r = 0.;
            rvr[q] = r;
        }
        // ---- z-momentum on z-faces (rhs is zero on the bottom and top faces) ----
        if (i <= nx - 1 && j <= ny - 1) {
            if (k == 0 || k == nz) { rwr[q] = 0.0; return; }
            const int hz = nz, lz = 0;
            double src;
            if (terr) {
                src = TA.zf(i, j, k, q);
            } else if (c2c2) {
                double xflux_hi = 0.25 * (ru[q + 1] + ru[q + 1 - sz]) * 1.0 * (w[q + 1] + w[q]);
                double xflux_lo = 0.25 * (ru[q] + ru[q - sz]) * 1.0 * (w[q - 1] + w[q]);
                double yflux_hi = 0.25 * (rv[q + sy] + rv[q + sy - sz]) * 1.0 * (w[q + sy] + w[q]);
                double yflux_lo = 0.25 * (rv[q] + rv[q - sz]) * 1.0 * (w[q - sy] + w[q]);
                double zflux_lo = 0.25 * (om[q] + om[q - sz]) * (w[q] + w[q - sz]);
                double zflux_hi = (k == hz) ? om[q] * w[q] : 0.25 * (om[q] + om[q + sz]) * (w[q] + w[q + sz]);
                src = (xflux_hi - xflux_lo) * dxi * (1.0 * 1.0) + (yflux_hi - yflux_lo) * dyi * (1.0 * 1.0) +
                      (zflux_hi - zflux_lo) * dzi;
            } else {
                double a_hi = 0.5 * (ru[q + 1] + ru[q + 1 - sz]) * 1.0;
                double xflux_hi = a_hi * interp_pick(mixed, ht, w, q + 1, 1, a_hi, hf);
                double a_lo = 0.5 * (ru[q] + ru[q - sz]) * 1.0;
                double xflux_lo = a_lo * interp_pick(mixed, ht, w, q, 1, a_lo, hf);
                double b_hi = 0.5 * (rv[q + sy] + rv[q + sy - sz]) * 1.0;
                double yflux_hi = b_hi * interp_pick(mixed, ht, w, q + sy, sy, b_hi, hf);
                double b_lo = 0.5 * (rv[q] + rv[q - sz]) * 1.0;
                double yflux_lo = b_lo * interp_pick(mixed, ht, w, q, sy, b_lo, hf);
                double zflux_hi, zflux_lo;
                if (k == hz) {
                    zflux_hi = om[q] * w[q];
                } else {
                    double c_hi = 0.5 * (om[q] + om[q + sz]);
                    double ih;
                    if (k == hz - 1) ih = interp_z_all(w, q + sz, sz, c_hi, 0, ERFX_ADV_CENTERED2);
                    else if (k == hz - 2 || k == lz + 1) {
                        if (vt != ERFX_ADV_CENTERED2 && vt != ERFX_ADV_UPWIND3)
                            ih = interp_z_all(w, q + sz, sz, c_hi, 0, ERFX_ADV_CENTERED4);
                        else
                            ih = interp_z_all(w, q + sz, sz, c_hi, vf, vt);
                    } else {
                        ih = interp_pick(mixed, vt, w, q + sz, sz, c_hi, vf);
                    }
                    zflux_hi = c_hi * ih;
                }
                if (k == lz) {
                    zflux_lo = om[q] * w[q];
                } else {
                    double c_lo = 0.5 * (om[q] + om[q - sz]);
                    double il;
                    if (k == lz + 1) il = interp_z_all(w, q, sz, c_lo, 0, ERFX_ADV_CENTERED2);
                    else if (k == lz + 2 || k == hz - 1) {
                        if (vt != ERFX_ADV_CENTERED2 && vt != ERFX_ADV_UPWIND3)
                            il = interp_z_all(w, q, sz, c_lo, 0, ERFX_ADV_CENTERED4);
                        else
                            il = interp_z_all(w, q, sz, c_lo, vf, vt);
                    } else {
                        il = interp_pick(mixed, vt, w, q, sz, c_lo, vf);
                    }
                    zflux_lo = c_lo * il;
                }
                double mfsq = 1.0 * 1.0;
                src = (xflux_hi - xflux_lo) * dxi * mfsq + (yflux_hi - yflux_lo) * dyi * mfsq +
                      (zflux_hi - zflux_lo) * dzi;
            }
            double r = -src;
            if (use_diff && terr) {   // :70-78, tau31 and tau32 in place of tau13 and tau23
                double diffContrib = ((t31[q + 1] - t31[q]) * dxi * 1.0 + (t32[q + sy] - t32[q]) * dyi * 1.0 +
                                      (t33[q] - t33[q - sz]) * dzi);
                diffContrib /= 0.5 * (Jd[q] + Jd[q - sz]);
                r -= diffContrib;
            } else if (use_diff)
                r -= ((t13[q + 1] - t13[q]) * dxi * 1.0 + (t23[q + sy] - t23[q]) * dyi * 1.0 +
                      (t33[q] - t33[q - sz]) * dzi);
            double zs;
            if (!moist) {
                zs = grav_z * 0.5 * ((cons[q] - r0[q]) + (cons[q - sz] - r0[q - sz]));
            } else if (btype == 1) {
                // ERF_make_buoyancy.cpp:195-208
                const double* c1 = cons + ERFX_RHOQ1 * ne;
                const double* c2 = cons + ERFX_RHOQ2 * ne;
                double rhop_hi = cons[q] + c1[q] + c2[q] - r0[q];
                double rhop_lo = cons[q - sz] + c1[q - sz] + c2[q - sz] - r0[q - sz];
                for (int q_offset = 2; q_offset < nqs; ++q_offset) {
                    rhop_hi += cons[q + (ERFX_RHOQ1 + q_offset) * ne];
                    rhop_lo += cons[q - sz + (ERFX_RHOQ1 + q_offset) * ne];
                }
                zs = grav_z * 0.5 * (rhop_hi + rhop_lo);
            } else {
                // ERF_make_buoyancy.cpp:272-323 (types 2, 4) and :347-372 (type 3)
                double qplus = moist_buoyancy_dev(btype, nqs, cons, pq1, pq2, pq3, pavg, nz, ne, q, k);
                double qminus = moist_buoyancy_dev(btype, nqs, cons, pq1, pq2, pq3, pavg, nz, ne, q - sz, k - 1);
                double qavg = 0.5 * (qplus + qminus);
                double r0avg = 0.5 * (pavg[k] + pavg[k - 1]);
                zs = -qavg * r0avg * grav_z;
            }
            if (cor) {
                double rho_u_loc = 0.25 * (ru[q + 1] + ru[q] + ru[q + 1 - sz] + ru[q - sz]);
                zs += cf * rho_u_loc * cosphi;
            }
            if (rayW && RA.act_nd(q, k)) {
                double rho_on_w_face = 0.5 * (cons[q] + cons[q - sz]);
                double ww = rw[q] / rho_on_w_face;
                zs -= RA.fac_nd(q, k) * (ww - ray[RAY_WBAR * n1 + k]) * rho_on_w_face;
            }
            zs += 0.5 * (cons[q] + cons[q - sz]) * g2;
            double met_h_zeta = terr ? met::h_zeta_Kface(zt, i, j, k, dzi) : 1;
            double gpz = dzi * (pp[q] - pp[q - sz]) / met_h_zeta;
            double qf = moist ? face_q(pq1, pq2, q, q - sz) : 0.0;
            r += (zs - gpz - apg2) / (1.0 + qf);
            rwr[q] = r;
        }
    };
    auto go = [&](auto tc) {
        launch_box_z(Box(0, 0, 0, nx, ny, nz), s, [=] __device__(int i, int j, int k) { body(i, j, k, tc); });
    };
    if (!c2c2 && ht == vt && ht == ERFX_ADV_UPWIND3) go(std::integral_constant<int, ERFX_ADV_UPWIND3>{});
    else if (!c2c2 && ht == vt && ht == ERFX_ADV_UPWIND5) go(std::integral_constant<int, ERFX_ADV_UPWIND5>{});
    else go(std::integral_constant<int, -1>{});
}

// ---- k-marching launch that evaluates every face once --------------------------
// A flux-form update needs, per cell, the fluxes through its six faces; a
// kernel with one lane per cell evaluates every interior face twice (as the
// hi face of one cell and the lo face of the next), and with WENO5 the face
// evaluation is most of the work.  Here a work-group covers a tile of 64
// cells in x, MR rows in y and up to mk levels in z: lane (lx, ly < MR) owns
// cell (i0 + lx, j0 + ly) and marches up through the levels; an extra wave
// (ly = MR) evaluates the lo y-faces of row j0 + MR and, in MR of its lanes,
// the x-face i0 + 64 of each row.  Per level:
//   1. the advected fields' values on the level, rows j0-3 .. j0+MR+2 and
//      columns i0-3 .. i0+66 (the tile and the reach of a six-point stencil),
//      are in LDS (`pl`): each x- and y-face takes its six stencil values from
//      there, each row of the level is read from memory once;
//   2. the z-stencils come from a ring of six levels per field in the owning
//      lane's registers, one new level loaded per level;
//   3. the lo x- and y-faces go to LDS (`xs`, `ys`), the hi z-face stays in
//      registers (the lo z-face is the previous level's hi face); after a
//      barrier each cell reads its hi x- and y-faces from LDS and is updated.
// The next level's rows are loaded into registers before the cell updates
// and stored to LDS after them, so their latency overlaps the updates.  A
// face's value is computed by the same expression from the same operands as
// in the per-cell kernel, so results are identical.
// Rows of cells per work-group: 15 in both forms (16 waves, the largest
// work-group; measured on the 512 x 512 x 96 moist terrain case, one WENO5
// scalar per launch: FP64 874 us at 15 rows, 1015 us at 12, 1154 us at 8,
// 1279 us at 4; mixed 717 us at 15 rows, 957 us at 9, 854 us at 4).  The
// registers (106 VGPRs FP64, 90 mixed) allow one 16-wave work-group per CU,
// so taller tiles mean more waves per CU as well as less y-halo.
template <bool mixed>
constexpr int march_rows() { return 15; }
constexpr int MARCH_MK = 12;   // levels per work-group (one WENO5 scalar at 15 rows: 882 / 879 / 859 / 879 / 925 / 971 us at 8 / 10 / 12 / 16 / 24 / 32)
constexpr int MARCH_PW = 72;   // LDS row of the level tile: columns i0-3 .. i0+66, padded

template <class Op, int MR>
__global__ void __launch_bounds__(64 * (MR + 1))
march_kernel(const Op op, int nx, int ny, int nz, unsigned gx, unsigned gy, int mk) {
    constexpr int NF = Op::NF;
    constexpr int PR = MR + 6;                        // rows of the level tile
    constexpr int RPT = (PR + MR) / (MR + 1);         // tile rows loaded per lane (rows ly, ly + MR + 1, ...)
    __shared__ double pl[NF][PR][MARCH_PW];
    __shared__ double xs[NF][MR][65];
    __shared__ double ys[NF][MR + 1][64];
    const Lay L = op.L;
    const unsigned t = xcd_remap(blockIdx.x, gridDim.x);
    const unsigned bx = t % gx, by = (t / gx) % gy, bz = t / (gx * gy);
    const int lx = threadIdx.x, ly = threadIdx.y;
    const int i0 = (int)bx * 64, j0 = (int)by * MR;
    const int i = i0 + lx;
    const int k0 = (int)bz * mk;
    const int k1 = min(k0 + mk, nz);
    const bool row = ly < MR;
    const int j = j0 + ly;                    // cell row (row lanes) or y-face row j0 + MR
    const bool cellok = row && i < nx && j < ny;
    const int ng = L.ng;
    // the level tile's loads of this lane: main column i0-3+lx of its rows and,
    // in lanes 0..5, the halo column i0+61+lx; points outside the layout read 0
    // (no face of the box uses them)
    const int cm = i0 - 3 + lx, ch = i0 + 61 + lx;
    const bool cm_ok = cm >= -ng && cm <= nx + ng, ch_ok = lx < 6 && ch <= nx + ng;
    double rm[NF][RPT], rh[NF][RPT];
    auto load_level = [&](int k) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int rr = ly + r * (MR + 1);
            const int jj = j0 - 3 + rr;
            const bool rok = rr < PR && jj >= -ng && jj <= ny + ng;
#pragma unroll
            for (int m = 0; m < NF; ++m) {
                const double* ps = op.field(m);
                rm[m][r] = (rok && cm_ok) ? ps[L.ix(cm, jj, k)] : 0.0;
                rh[m][r] = (rok && ch_ok) ? ps[L.ix(ch, jj, k)] : 0.0;
            }
        }
    };
    auto store_level = [&]() {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int rr = ly + r * (MR + 1);
            if (rr >= PR) break;
#pragma unroll
            for (int m = 0; m < NF; ++m) {
                pl[m][rr][lx] = rm[m][r];
                if (lx < 6) pl[m][rr][64 + lx] = rh[m][r];
            }
        }
    };
    // z ring of the owning lane: zr[m][0..5] = levels k-2 .. k+3 at level k
    // (levels -3 .. nz+2 lie in the layout: at least 3 ghost layers)
    double zr[NF][6], zn[NF], gzl[NF], gzh[NF];
#pragma unroll
    for (int m = 0; m < NF; ++m) {
        zn[m] = 0.0; gzl[m] = 0.0; gzh[m] = 0.0;
#pragma unroll
        for (int d = 0; d < 6; ++d) zr[m][d] = 0.0;
    }
    const long qc = cellok ? L.ix(i, j, 0) : 0;
    const long sz = L.sz;
    if (cellok) {
#pragma unroll
        for (int m = 0; m < NF; ++m) {
            const double* ps = op.field(m);
            const double lo3 = ps[qc + (long)(k0 - 3) * sz];
#pragma unroll
            for (int d = 0; d < 6; ++d) zr[m][d] = ps[qc + (long)(k0 - 2 + d) * sz];
            // lo face of level k0: stencil levels k0-3 .. k0+2
            gzl[m] = op.face_val(op.az, L.ix(i, j, k0), op.vt, op.vf, zr[m][4], zr[m][3], zr[m][2], zr[m][1],
                                 zr[m][0], lo3);
        }
    }
    load_level(k0);
    store_level();
    for (int k = k0; k < k1; ++k) {
        __syncthreads();   // the level's tile is in LDS; last level's cell updates are done
        if (row) {
            // the lo x- and y-faces of this cell and the hi z-face, evaluated
            // for every lane (a face outside the box gets a zero flux and is
            // discarded) so that the three evaluations form one block of code
            const int c = lx + 3, rr = ly + 3;   // this cell's place in the tile
            const bool okx = i <= nx && j < ny, oky = i < nx && j <= ny;
            const long q = L.ix(i, j, k);
            const double fx = okx ? op.ax[q] : 0.0;
            const double fy = oky ? op.ay[q] : 0.0;
            const double fz = cellok ? op.az[q + L.sz] : 0.0;
#pragma unroll
            for (int m = 0; m < NF; ++m) {
                const double vx = op.face_of(fx, op.ht, op.hf, pl[m][rr][c + 2], pl[m][rr][c + 1], pl[m][rr][c],
                                             pl[m][rr][c - 1], pl[m][rr][c - 2], pl[m][rr][c - 3]);
                const double vy = op.face_of(fy, op.ht, op.hf, pl[m][rr + 2][c], pl[m][rr + 1][c], pl[m][rr][c],
                                             pl[m][rr - 1][c], pl[m][rr - 2][c], pl[m][rr - 3][c]);
                const double vz = op.face_of(fz, op.vt, op.vf, zr[m][5], zr[m][4], zr[m][3], zr[m][2], zr[m][1],
                                             zr[m][0]);
                xs[m][ly][lx] = okx ? vx : 0.0;
                ys[m][ly][lx] = oky ? vy : 0.0;
                gzh[m] = vz;
            }
        } else {
            // the y-faces of row j0 + MR and, in lanes 0 .. MR-1, x-face i0 + 64 of row j0 + lx
            const int c = lx + 3, rr = MR + 3;
            const bool oky = i < nx && j <= ny;
            const int ix = i0 + 64, jr = j0 + (lx < MR ? lx : 0), r2 = (lx < MR ? lx : 0) + 3, c2 = 64 + 3;
            const bool okx = lx < MR && ix <= nx && jr < ny;
            const double fy = oky ? op.ay[L.ix(i, j, k)] : 0.0;
            const double fx = okx ? op.ax[L.ix(ix, jr, k)] : 0.0;
#pragma unroll
            for (int m = 0; m < NF; ++m) {
                const double vy = op.face_of(fy, op.ht, op.hf, pl[m][rr + 2][c], pl[m][rr + 1][c], pl[m][rr][c],
                                             pl[m][rr - 1][c], pl[m][rr - 2][c], pl[m][rr - 3][c]);
                const double vx = op.face_of(fx, op.ht, op.hf, pl[m][r2][c2 + 2], pl[m][r2][c2 + 1], pl[m][r2][c2],
                                             pl[m][r2][c2 - 1], pl[m][r2][c2 - 2], pl[m][r2][c2 - 3]);
                ys[m][MR][lx] = oky ? vy : 0.0;
                if (lx < MR) xs[m][lx][64] = okx ? vx : 0.0;
            }
        }
        __syncthreads();   // faces in LDS; every read of the level tile is done
        const bool more = k + 1 < k1;
        if (more) {
            load_level(k + 1);
            if (cellok) {
#pragma unroll
                for (int m = 0; m < NF; ++m) zn[m] = op.field(m)[qc + (long)(k + 4) * sz];
            }
        }
        if (cellok) {
            double gx0[NF], gx1[NF], gy0[NF], gy1[NF];
#pragma unroll
            for (int m = 0; m < NF; ++m) {
                gx0[m] = xs[m][ly][lx];
                gx1[m] = xs[m][ly][lx + 1];
                gy0[m] = ys[m][ly][lx];
                gy1[m] = ys[m][ly + 1][lx];
            }
            op.cell(i, j, k, gx0, gx1, gy0, gy1, gzl, gzh);
#pragma unroll
            for (int m = 0; m < NF; ++m) gzl[m] = gzh[m];
        }
        if (more) {
            // every read of this level's tile was before the barrier above, and
            // the next level's faces overwrite xs / ys only after the barrier at
            // the loop's top, which every lane reaches after its cell update
            store_level();
            if (cellok) {
#pragma unroll
                for (int m = 0; m < NF; ++m) {
#pragma unroll
                    for (int d = 0; d < 5; ++d) zr[m][d] = zr[m][d + 1];
                    zr[m][5] = zn[m];
                }
            }
        }
    }
}

template <class Op, int MR>
inline void launch_march_rows(const Op& op, int nx, int ny, int nz, int mk, hipStream_t s) {
    const unsigned gx = (unsigned)((nx + 63) / 64), gy = (unsigned)((ny + MR - 1) / MR);
    const unsigned gz = (unsigned)((nz + mk - 1) / mk);
    hipLaunchKernelGGL((march_kernel<Op, MR>), dim3(gx * gy * gz), dim3(64, MR + 1, 1), 0, s, op, nx, ny, nz, gx, gy,
                       mk);
}
// A box whose default tiling gives fewer work-groups than the GPU has CUs
// (the 64^3 rising thermal: 24) takes shorter stacks of levels and, if that
// is not enough, 4-row tiles: a small box's march is bound by the latency of
// its levels in sequence, not by the extra halo reads.  The tiling does not
// change any value.
constexpr int MARCH_MIN_GROUPS = 256;
template <class Op>
inline void launch_march(const Op& op, int nx, int ny, int nz, hipStream_t s) {
    constexpr int MR = march_rows<Op::MIXED>();
    auto groups = [&](int mr, int mk) {
        return ((nx + 63) / 64) * ((ny + mr - 1) / mr) * ((nz + mk - 1) / mk);
    };
    int mk = MARCH_MK;
    while (mk > 4 && groups(MR, mk) < MARCH_MIN_GROUPS) mk = mk / 2 > 4 ? mk / 2 : 4;
    if (MR != 4 && groups(MR, mk) < MARCH_MIN_GROUPS) launch_march_rows<Op, 4>(op, nx, ny, nz, mk, s);
    else launch_march_rows<Op, MR>(op, nx, ny, nz, mk, s);
}

// The scalar update of erf_slow_rhs_post as a marching operator: NF scalars
// of one block, face values as in post_block below.
template <bool mixed, int NF_, int TT = -1>
struct PostOp {
    static constexpr int NF = NF_;
    static constexpr bool MIXED = mixed;
    Lay L;
    int nx, ny;
    double dxi, dyi, dzi;
    const double* Jc;
    int ht, vt;
    double hf, vf;
    bool c2c2, use_diff, consA, turb, clip, copy_slow;
    double ae, slow_dt;
    bool dxlo, dxhi, dylo, dyhi;
    const double* mu_h;
    const double* mu_v;
    int n0, ncons;
    const double *ax, *ay, *az;
    double* cur;
    double* nc;
    const double* oc;
    double* crb;
    const double* prim;
    const double* dstate;     // terrain grid: the diffusion terms (DiffusionSrcForState_T), else null
    // the advected field m: prim component n0 + m - 1
    __device__ __forceinline__ const double* field(int m) const { return prim + (long)(n0 + m - 1) * L.ne; }
    // flux through face q with mass flux a[q] and stencil values sm3..sp2
    __device__ __forceinline__ double face_val(const double* a, long q, int t, double fr, double sp2, double sp1,
                                               double s, double sm1, double sm2, double sm3) const {
        return face_of(a[q], t, fr, sp2, sp1, s, sm1, sm2, sm3);
    }
    // the same with the mass flux f at hand
    __device__ __forceinline__ double face_of(double f, int t, double fr, double sp2, double sp1, double s, double sm1,
                                              double sm2, double sm3) const {
        if constexpr (TT >= 0) return f * interp_scalar_vals_t<mixed, TT>(t, sp2, sp1, s, sm1, sm2, sm3, f, fr);
        return c2c2 ? f * (0.5 * (s + sm1)) : f * interp_scalar_vals<mixed>(t, sp2, sp1, s, sm1, sm2, sm3, f, fr);
    }
    __device__ __forceinline__ void cell(int i, int j, int k, const double* gx0, const double* gx1, const double* gy0,
                                         const double* gy1, const double* gz0, const double* gz1) const {
        const long q = L.ix(i, j, k);
        const long ne = L.ne, sy = L.sy, sz = L.sz;
        if (copy_slow)
            for (int c = 2; c < ncons; ++c) cur[q + c * ne] = nc[q + c * ne];
        for (int m = 0; m < NF; ++m) {
            const int n = n0 + m;
            const double* ps = prim + (n - 1) * ne;
            double invdetJ = Jc ? ((Jc[q] > 0.) ? 1. / Jc[q] : 1.) : 1.;
            double r = -invdetJ * (1.0 * 1.0) * ((gx1[m] - gx0[m]) * dxi + (gy1[m] - gy0[m]) * dyi + (gz1[m] - gz0[m]) * dzi);
            if (use_diff && dstate) {
                r -= dstate[q + n * ne];
            } else if (use_diff) {
                const int dsx0 = (dxlo && i == 0) ? -1 : 0, dsx1 = (dxhi && i == nx - 1) ? 1 : 0;
                const int dsy0 = (dylo && j == 0) ? -1 : 0, dsy1 = (dyhi && j == ny - 1) ? 1 : 0;
                double dfx0 = diff_flux_bc(nc, ps, mu_h, q, 1, dxi, ae, consA, turb, dsx0) * 1.0;
                double dfx1 = diff_flux_bc(nc, ps, mu_h, q + 1, 1, dxi, ae, consA, turb, dsx1) * 1.0;
                double dfy0 = diff_flux_bc(nc, ps, mu_h, q, sy, dyi, ae, consA, turb, dsy0) * 1.0;
                double dfy1 = diff_flux_bc(nc, ps, mu_h, q + sy, sy, dyi, ae, consA, turb, dsy1) * 1.0;
                double dfz0 = diff_flux(nc, ps, mu_v, q, sz, dzi, ae, consA, turb);
                double dfz1 = diff_flux(nc, ps, mu_v, q + sz, sz, dzi, ae, consA, turb);
                r -= (dfx1 - dfx0) * dxi * 1.0 + (dfy1 - dfy0) * dyi * 1.0 + (dfz1 - dfz0) * dzi;
            }
            crb[q + n * ne] = r;
            double v = oc[q + n * ne] + slow_dt * r;
            if (clip) v = (v < 0.0) ? 0.0 : v;   // amrex::max(v, 0.0)
            cur[q + n * ne] = v;
            nc[q + n * ne] = v;   // S_new <- S_data for this component (nothing reads S_new's scalars here)
        }
    }
};

// One block of slow scalars of erf_slow_rhs_post (ERF_slow_rhs_post.cpp:364-497):
// advection with the time-averaged momenta (AdvectionSrcForScalars), diffusion
// (DiffusionSrcForState_N), then cur = old + dt * rhs, clipped at 0 for the
// moisture species.  copy_slow: first copy every slow comp of S_new into cur.
struct ScalarBlock {
    int n0, ncomp;            // cons comps [n0, n0 + ncomp)
    int ht, vt;               // advection types (a WENO horizontal type also in z)
    double hf, vf;
    int mu_h, mu_v;           // EddyDiff comps
    bool clip;
};
static void post_block(const Params& P, const SlowWork& W, const ScalarBlock& B, bool copy_slow, double slow_dt,
                       const erfx_state& rhs, const erfx_state& old, const erfx_state& snew, const erfx_state& data,
                       const double* prim, const erfx_state& scratch, const double* eddy, hipStream_t s) {
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne, sy = L.sy, sz = L.sz;
    const double dxi = P.dxInv[0], dyi = P.dxInv[1], dzi = P.dxInv[2];
    const int ncons = P.cfg.ncons;
    const double* Jc = P.terrain ? P.detJ : nullptr;   // 1/detJ on a terrain grid (:298, :361)
    const int ht = B.ht;
    const int vt = is_weno(ht) ? ht : B.vt;
    const double hf = B.hf, vf = B.vf;
    const bool c2c2 = (ht == ERFX_ADV_CENTERED2 && vt == ERFX_ADV_CENTERED2);
    const bool mixed_rt = P.cfg.precision == ERFX_PREC_MIXED;
    const bool use_diff = P.use_diff, consA = P.consA, turb = P.use_turb;
    const double ae = consA ? P.cfg.alpha_C : P.cfg.rho0_trans * P.cfg.alpha_C;
    // Dirichlet scalar faces (Inflow): one-sided diffusive fluxes there
    const bool dxlo = P.bck[3][0] == EXT_DIR, dxhi = P.bck[3][3] == EXT_DIR;
    const bool dylo = P.bck[3][1] == EXT_DIR, dyhi = P.bck[3][4] == EXT_DIR;
    const double* mu_h = eddy ? eddy + B.mu_h * ne : nullptr;
    const double* mu_v = eddy ? eddy + B.mu_v * ne : nullptr;
    const int n0 = B.n0, nc_b = B.ncomp;
    const bool clip = B.clip;
    const double* ax = scratch.xmom;
    const double* ay = scratch.ymom;
    const double* az = scratch.zmom;
    double* cur = data.cons;
    double* nc = snew.cons;
    const double* oc = old.cons;
    double* crb = rhs.cons;
    // on a terrain grid the diffusion terms are made beforehand (DiffusionSrcForState_T)
    const double* dst = nullptr;
    if (use_diff && P.terrain) {
        diffusion_state_T(P, W, n0, nc_b, nc, prim, s);
        dst = W.dstate;
    }
    if (!g_rec) {
        // one evaluation per face (the batched ensembles record the per-cell
        // form below), one launch per scalar of the block: the six-level z
        // rings and the staged rows of several fields at once would exceed the
        // registers of four waves per SIMD.  The first launch makes the copy of
        // the slow components; the later ones read only prim, the face mass
        // fluxes and the old state, which no launch of the block writes.
        auto run = [&](auto mc, auto tc, int m) {
            constexpr bool mixed = decltype(mc)::value;
            PostOp<mixed, 1, decltype(tc)::value> op;
            op.L = L; op.nx = nx; op.ny = ny;
            op.dxi = dxi; op.dyi = dyi; op.dzi = dzi; op.Jc = Jc;
            op.ht = ht; op.vt = vt; op.hf = hf; op.vf = vf;
            op.c2c2 = c2c2; op.use_diff = use_diff; op.consA = consA; op.turb = turb; op.clip = clip;
            op.copy_slow = copy_slow && m == 0; op.ae = ae; op.slow_dt = slow_dt;
            op.dxlo = dxlo; op.dxhi = dxhi; op.dylo = dylo; op.dyhi = dyhi;
            op.mu_h = mu_h; op.mu_v = mu_v; op.n0 = n0 + m; op.ncons = ncons;
            op.ax = ax; op.ay = ay; op.az = az; op.cur = cur; op.nc = nc; op.oc = oc; op.crb = crb; op.prim = prim;
            op.dstate = dst;
            launch_march(op, nx, ny, nz, s);
        };
        // one type in all three directions (WENO5, WENO-Z5, U3, U5): that type
        // fixed at compile time
        const int fixed = (!c2c2 && ht == vt && (ht == ERFX_ADV_WENO5 || ht == ERFX_ADV_WENOZ5 ||
                                                 ht == ERFX_ADV_UPWIND3 || ht == ERFX_ADV_UPWIND5)) ? ht : -1;
        auto typed = [&](auto mc, int m) {
            switch (fixed) {
                case ERFX_ADV_WENO5: run(mc, std::integral_constant<int, ERFX_ADV_WENO5>{}, m); break;
                case ERFX_ADV_WENOZ5: run(mc, std::integral_constant<int, ERFX_ADV_WENOZ5>{}, m); break;
                case ERFX_ADV_UPWIND3: run(mc, std::integral_constant<int, ERFX_ADV_UPWIND3>{}, m); break;
                case ERFX_ADV_UPWIND5: run(mc, std::integral_constant<int, ERFX_ADV_UPWIND5>{}, m); break;
                default: run(mc, std::integral_constant<int, -1>{}, m); break;
            }
        };
        for (int m = 0; m < nc_b; ++m) {
            if (mixed_rt) typed(std::true_type{}, m);
            else typed(std::false_type{}, m);
        }
        return;
    }
    launch_box_prec(mixed_rt, Box(0, 0, 0, nx - 1, ny - 1, nz - 1), s, [=] __device__(int i, int j, int k, auto mc) {
        constexpr bool mixed = decltype(mc)::value;
        const long q = L.ix(i, j, k);
        if (copy_slow)
            for (int c = 2; c < ncons; ++c) cur[q + c * ne] = nc[q + c * ne];
        double fx0 = ax[q], fx1 = ax[q + 1], fy0 = ay[q], fy1 = ay[q + sy], fz0 = az[q], fz1 = az[q + sz];
        for (int m = 0; m < nc_b; ++m) {
            const int n = n0 + m;
            const double* ps = prim + (n - 1) * ne;
            double gx0, gx1, gy0, gy1, gz0, gz1;
            if (c2c2) {
                gx0 = fx0 * (0.5 * (ps[q] + ps[q - 1]));
                gx1 = fx1 * (0.5 * (ps[q + 1] + ps[q]));
                gy0 = fy0 * (0.5 * (ps[q] + ps[q - sy]));
                gy1 = fy1 * (0.5 * (ps[q + sy] + ps[q]));
                gz0 = fz0 * (0.5 * (ps[q] + ps[q - sz]));
                gz1 = fz1 * (0.5 * (ps[q + sz] + ps[q]));
            } else {
                gx0 = fx0 * interp_scalar_pick(mixed, ht, ps, q, 1, fx0, hf);
                gx1 = fx1 * interp_scalar_pick(mixed, ht, ps, q + 1, 1, fx1, hf);
                gy0 = fy0 * interp_scalar_pick(mixed, ht, ps, q, sy, fy0, hf);
                gy1 = fy1 * interp_scalar_pick(mixed, ht, ps, q + sy, sy, fy1, hf);
                gz0 = fz0 * interp_scalar_pick(mixed, vt, ps, q, sz, fz0, vf);
                gz1 = fz1 * interp_scalar_pick(mixed, vt, ps, q + sz, sz, fz1, vf);
            }
            double invdetJ = Jc ? ((Jc[q] > 0.) ? 1. / Jc[q] : 1.) : 1.;
            double r = -invdetJ * (1.0 * 1.0) * ((gx1 - gx0) * dxi + (gy1 - gy0) * dyi + (gz1 - gz0) * dzi);
            if (use_diff && dst) {
                r -= dst[q + n * ne];
            } else if (use_diff) {
                const int dsx0 = (dxlo && i == 0) ? -1 : 0, dsx1 = (dxhi && i == nx - 1) ? 1 : 0;
                const int dsy0 = (dylo && j == 0) ? -1 : 0, dsy1 = (dyhi && j == ny - 1) ? 1 : 0;
                double dfx0 = diff_flux_bc(nc, ps, mu_h, q, 1, dxi, ae, consA, turb, dsx0) * 1.0;
                double dfx1 = diff_flux_bc(nc, ps, mu_h, q + 1, 1, dxi, ae, consA, turb, dsx1) * 1.0;
                double dfy0 = diff_flux_bc(nc, ps, mu_h, q, sy, dyi, ae, consA, turb, dsy0) * 1.0;
                double dfy1 = diff_flux_bc(nc, ps, mu_h, q + sy, sy, dyi, ae, consA, turb, dsy1) * 1.0;
                double dfz0 = diff_flux(nc, ps, mu_v, q, sz, dzi, ae, consA, turb);
                double dfz1 = diff_flux(nc, ps, mu_v, q + sz, sz, dzi, ae, consA, turb);
                r -= (dfx1 - dfx0) * dxi * 1.0 + (dfy1 - dfy0) * dyi * 1.0 + (dfz1 - dfz0) * dzi;
            }
            crb[q + n * ne] = r;
            double v = oc[q + n * ne] + slow_dt * r;
            if (clip) v = (v < 0.0) ? 0.0 : v;   // amrex::max(v, 0.0)
            cur[q + n * ne] = v;
            nc[q + n * ne] = v;   // S_new <- S_data for this component (nothing reads S_new's scalars here)
        }
    });
}

// erf_slow_rhs_post (Source/TimeIntegration/ERF_slow_rhs_post.cpp:49-559): the
// slow variables are RhoScalar (:168-172) and, with moisture, the n_qstate
// species as one block with the moist advection types (:173-175, :364-434);
// then S_new <- S_data.
void slow_rhs_post(const Params& P, const SlowWork& W, int nrk, double slow_dt, const erfx_state& rhs,
                   const erfx_state& old, const erfx_state& snew, const erfx_state& data, const double* prim,
                   const erfx_state& scratch, const double* eddy, hipStream_t s) {
    (void)nrk;
    const Lay L = P.L;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const long ne = L.ne;
    const int ncons = P.cfg.ncons;
    const erfx_config& c = P.cfg;
    ScalarBlock scal{ERFX_RHOSCALAR, 1, c.dryscal_horiz_adv, c.dryscal_vert_adv, c.dryscal_horiz_upw_frac,
                     c.dryscal_vert_upw_frac, 4 /* Scalar_h */, 10 /* Scalar_v */, false};
    post_block(P, W, scal, true, slow_dt, rhs, old, snew, data, prim, scratch, eddy, s);
    if (P.moist) {
        ScalarBlock moist{ERFX_RHOQ1, P.n_qstate, c.moistscal_horiz_adv, c.moistscal_vert_adv,
                          c.moistscal_horiz_upw_frac, c.moistscal_vert_upw_frac, 5 /* Q_h */, 11 /* Q_v */, true};
        post_block(P, W, moist, false, slow_dt, rhs, old, snew, data, prim, scratch, eddy, s);
    }
    double* cur = data.cons;
    double* nc = snew.cons;
    // S_new <- S_data after every cell is updated (neighbours read S_new's rho
    // above).  The scalar blocks wrote their components to both, and with
    // copy_slow the other slow components went from S_new to S_data, so rho and
    // rho*theta are what is left of cons.
    launch_box(Box(0, 0, 0, nx, ny, nz), s, [=] __device__(int i, int j, int k) {
        const long q = L.ix(i, j, k);
        if (i <= nx - 1 && j <= ny - 1 && k <= nz - 1)
            for (int c = 0; c < 2; ++c) nc[q + c * ne] = cur[q + c * ne];
        if (j <= ny - 1 && k <= nz - 1) snew.xmom[q] = data.xmom[q];
        if (i <= nx - 1 && k <= nz - 1) snew.ymom[q] = data.ymom[q];
        if (i <= nx - 1 && j <= ny - 1) snew.zmom[q] = data.zmom[q];
    });
}

}  // namespace erfx
