// The line-vortex kernel K(line, p) of the reference (src/core/flow_solver.c:114-135):
//
//   dr1 = p - r1, dr2 = p - r2, t_i = e . dr_i, h^2 = ((|dr1|^2 - t1^2) + (|dr2|^2 - t2^2)) / 2
//   K = 0                                            if |d| < tol or h^2 < tol^2
//   K = cross(dr1, e) * (atan2(t2, h) - atan2(t1, h)) / h   otherwise
//
// |K| is the angle the segment subtends from p (no 1/(4 pi) factor).  The file must be
// compiled with -ffp-contract=off: the dot products, h^2 and the cut-off test below then
// round exactly as the reference's GCC x86-64 build (left to right, no FMA), so the cut-off
// branch is decided identically and the cancellation in h^2 near a segment's line (which
// the reference's values carry) is reproduced.  FMAs appear only where written explicitly:
// in the angle evaluation and, for the non-reference forms, in the cross product and the
// circulation-weighted accumulation (1-ulp differences, no effect on the cut-off).
//
// Angle forms:
//   PVL_FP64_REF  the reference's expression: two FP64 atan2 calls of the HIP math library.
//   PVL_FP64      one angle: atan2(t2, h) - atan2(t1, h) = atan2((t2 - t1) h, h^2 + t1 t2).
//                 Both sides lie in (-pi, pi): with r_i = sqrt(h^2 + t_i^2) > 0, the sine and
//                 cosine of the difference are (t2 - t1) h / (r1 r2) and (h^2 + t1 t2) /
//                 (r1 r2).  The error of this form is O(eps) relative to the angle everywhere,
//                 whereas the difference of two atan2 values loses digits in the far field
//                 (|angle| << pi); the two forms agree to that loss (~1e-14 at r/len ~ 100).
//                 atan2 is this library's own (pvl_atan2 below), 1/h comes from a
//                 Newton-refined reciprocal square root.
//   PVL_FP32      as PVL_FP64 with the angle, h and 1/h in FP32 (cut-off, dots, h^2 and
//                 cross(dr1, e) stay FP64); ~1e-7 relative.
#pragma once

#include "pvl_internal.hpp"

enum
{
    PVL_FP64_REF_FORM = 2
};

// atan(a) = a + a s P(s), s = a^2, 0 <= a <= 1.  Minimax coefficients (maximum relative
// error of the polynomial 2.8e-17 for FP64, 1.7e-8 for FP32) fitted by
// pyvl_amd/csrc/tools/fit_atan.py (Lawson iteration in 60-digit arithmetic).
__device__ __forceinline__ double atan_poly64(double a)
{
    const double s = a * a;
    double p = -2.0258217196429153806e-5;
    p = fma(p, s, 2.2301921440372192015e-4);
    p = fma(p, s, -1.1640578327914545487e-3);
    p = fma(p, s, 3.855937704218664755e-3);
    p = fma(p, s, -9.1844913587584364515e-3);
    p = fma(p, s, 1.6977946435648326745e-2);
    p = fma(p, s, -2.5826708868270838343e-2);
    p = fma(p, s, 3.4067745269857893501e-2);
    p = fma(p, s, -4.0926344607047341725e-2);
    p = fma(p, s, 4.6739479480120956234e-2);
    p = fma(p, s, -5.2392324387842287365e-2);
    p = fma(p, s, 5.877307626399871717e-2);
    p = fma(p, s, -6.6658603353628207695e-2);
    p = fma(p, s, 7.6922129266775500417e-2);
    p = fma(p, s, -9.0909012350179866105e-2);
    p = fma(p, s, 1.1111110678724637807e-1);
    p = fma(p, s, -1.4285714271333841475e-1);
    p = fma(p, s, 1.999999999975499914e-1);
    p = fma(p, s, -3.3333333333331858665e-1);
    return fma(a * s, p, a);
}

__device__ __forceinline__ float atan_poly32(float a)
{
    const float s = a * a;
    float p = 2.9206460455115316945e-3f;
    p = fmaf(p, s, -1.6367743921648625112e-2f);
    p = fmaf(p, s, 4.3211563153423941466e-2f);
    p = fmaf(p, s, -7.5521892793692703723e-2f);
    p = fmaf(p, s, 1.0665993049895164418e-1f);
    p = fmaf(p, s, -1.4211052413584158043e-1f);
    p = fmaf(p, s, 1.9993772488057382827e-1f);
    p = fmaf(p, s, -3.3333152721644904729e-1f);
    return fmaf(a * s, p, a);
}

// atan2(y, x) for finite arguments not both zero: a = min(|x|,|y|) / max(|x|,|y|) in [0, 1]
// (the hardware reciprocal, within 2^-24 relative on gfx950 (measured), one Newton step to
// ~2^-48, then the quotient corrected by its residual: within about an ulp), atan(a) from
// the polynomial, then the octant: pi/2 - . when |y| > |x|, pi - . when x < 0, sign of y.
__device__ __forceinline__ double pvl_atan2(double y, double x)
{
    const double ax = fabs(x), ay = fabs(y);
    const double mn = fmin(ax, ay), mx = fmax(ax, ay);
    double r = __builtin_amdgcn_rcp(mx);
    r = fma(r, fma(-mx, r, 1.0), r);
    double a = mn * r;
    a = fma(r, fma(-mx, a, mn), a);
    double t = atan_poly64(a);
    t = (ay > ax) ? (1.57079632679489661923 - t) : t;
    t = (x < 0.0) ? (3.14159265358979323846 - t) : t;
    return copysign(t, y);
}

__device__ __forceinline__ float pvl_atan2f(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mn = fminf(ax, ay), mx = fmaxf(ax, ay);
    const float a = mn * __builtin_amdgcn_rcpf(mx);
    float t = atan_poly32(a);
    t = (ay > ax) ? (1.57079632679489661923f - t) : t;
    t = (x < 0.0f) ? (3.14159265358979323846f - t) : t;
    return copysignf(t, y);
}

// 1/sqrt(x) for x > 0 normal: the hardware estimate (within 2^-23 relative on gfx950,
// measured) and two Newton steps y += y (1/2 - (x/2) y^2).
__device__ __forceinline__ double pvl_rsqrt(double x)
{
    const double hx = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = fma(y, fma(-(hx * y), y, 0.5), y);
    y = fma(y, fma(-(hx * y), y, 0.5), y);
    return y;
}

// Scalar factor of K (angle / h; the cut-off is decided here) and dr1 = p - r1.
template <int PREC>
__device__ __forceinline__ double line_m(const LineGeom &g, double px, double py, double pz, double tol, double tol2,
                                         double &d1x, double &d1y, double &d1z, bool &cut)
{
    d1x = px - g.r1x, d1y = py - g.r1y, d1z = pz - g.r1z;
    const double d2x = px - g.r2x, d2y = py - g.r2y, d2z = pz - g.r2z;
    const double t1 = g.ex * d1x + g.ey * d1y + g.ez * d1z;
    const double t2 = g.ex * d2x + g.ey * d2y + g.ez * d2z;
    const double n1 = (d1x * d1x + d1y * d1y + d1z * d1z) - t1 * t1;
    const double n2 = (d2x * d2x + d2y * d2y + d2z * d2z) - t2 * t2;
    const double h2 = (n1 + n2) / 2.0;
    double m;
    if constexpr (PREC == PVL_FP64_REF_FORM)
    {
        cut = (g.len < tol) || (h2 < tol2);
        const double hh = cut ? 1.0 : h2;
        const double h = sqrt(hh);
        m = (atan2(t2, h) - atan2(t1, h)) / h;
    }
    else
    {
        // Only h2 < tol^2 is tested here: a line shorter than tol has direction 0
        // (line_geom_kernel), so t1 = t2 = 0, the angle atan2(0, h2 > 0) is 0 and the line
        // adds exactly 0, as the reference's cut-off does.  A cut pair's h2 (possibly <= 0)
        // may give NaN or Inf below; the caller replaces the pair's value by 0.
        cut = h2 < tol2;
        if constexpr (PREC == PVL_FP64)
        {
            const double ih = pvl_rsqrt(h2);
            const double h = h2 * ih;
            const double dphi = pvl_atan2((t2 - t1) * h, fma(t1, t2, h2));
            m = dphi * ih;
        }
        else
        {
            const float hf = sqrtf((float)h2);
            const float ihf = __builtin_amdgcn_rcpf(hf);
            const float num = (float)(t2 - t1) * hf;
            const float den = (float)fma(t1, t2, h2);
            m = (double)(pvl_atan2f(num, den) * ihf);
        }
    }
    return m;
}

// K(line, p).  The reference form keeps the reference's cross product and products; the
// others fuse the cross product's second product into an FMA (the cut-off decision and
// every quantity it depends on are unchanged).
template <int PREC>
__device__ __forceinline__ void line_k(const LineGeom &g, double px, double py, double pz, double tol, double tol2,
                                       double &kx, double &ky, double &kz)
{
    double d1x, d1y, d1z;
    bool cut;
    const double m = line_m<PREC>(g, px, py, pz, tol, tol2, d1x, d1y, d1z, cut);
    if constexpr (PREC == PVL_FP64_REF_FORM)
    {
        const double cx = d1y * g.ez - d1z * g.ey;
        const double cy = d1z * g.ex - d1x * g.ez;
        const double cz = d1x * g.ey - d1y * g.ex;
        kx = cut ? 0.0 : cx * m;
        ky = cut ? 0.0 : cy * m;
        kz = cut ? 0.0 : cz * m;
    }
    else
    {
        const double cx = fma(d1y, g.ez, -(d1z * g.ey));
        const double cy = fma(d1z, g.ex, -(d1x * g.ez));
        const double cz = fma(d1x, g.ey, -(d1y * g.ex));
        kx = cut ? 0.0 : cx * m;
        ky = cut ? 0.0 : cy * m;
        kz = cut ? 0.0 : cz * m;
    }
}

// v += w K(line, p): the point-velocity inner step (w = the line's circulation).  The
// non-reference forms scale once (w m) and accumulate with FMAs.
template <int PREC>
__device__ __forceinline__ void line_k_acc(const LineGeom &g, double px, double py, double pz, double tol,
                                           double tol2, double w, double &vx, double &vy, double &vz)
{
    if constexpr (PREC == PVL_FP64_REF_FORM)
    {
        double kx, ky, kz;
        line_k<PREC>(g, px, py, pz, tol, tol2, kx, ky, kz);
        vx = vx + w * kx;
        vy = vy + w * ky;
        vz = vz + w * kz;
    }
    else
    {
        double d1x, d1y, d1z;
        bool cut;
        const double m = line_m<PREC>(g, px, py, pz, tol, tol2, d1x, d1y, d1z, cut);
        const double wm = cut ? 0.0 : w * m;
        vx = fma(wm, fma(d1y, g.ez, -(d1z * g.ey)), vx);
        vy = fma(wm, fma(d1z, g.ex, -(d1x * g.ez)), vy);
        vz = fma(wm, fma(d1x, g.ey, -(d1y * g.ex)), vz);
    }
}

