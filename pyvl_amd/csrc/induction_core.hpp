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
//                 (r1 r2).  The error of this form is a few 1e-15 relative to the angle
//                 everywhere (the atan polynomial, the quotient and 1/h below, each ~1e-15),
//                 whereas the difference of two atan2 values loses digits in the far field
//                 (|angle| << pi); the two forms agree to that loss (~1e-14 at r/len ~ 100).
//                 atan2 is this library's own (pvl_atan2 below), 1/h comes from a
//                 refined reciprocal square root.  The wake's self-induction takes h^2 from
//                 the cross product instead (cross_m below).
//   PVL_FP32      as PVL_FP64 with the angle, h and 1/h in FP32 (cut-off, dots, h^2 and
//                 cross(dr1, e) stay FP64); ~1e-7 relative.
#pragma once

#include "pvl_internal.hpp"

enum
{
    PVL_FP64_REF_FORM = 2
};

// atan(a) = a + a s P(s), s = a^2, 0 <= a <= 1.  Minimax coefficients fitted by
// pyvl_amd/csrc/tools/fit_atan.py (Lawson iteration in 60-digit arithmetic): P of degree 16
// for FP64 (maximum relative error of the polynomial 1.1e-15, ~5 ulp: the pair's other
// rounding errors are of that order), degree 7 for FP32 (1.7e-8).
// SGPR_COEF: the Horner steps as one inline-asm block of v_fma_f64 with the coefficients in
// SGPRs.  In kernels whose loop has enough control flow around the evaluation the compiler
// otherwise keeps the loop-invariant coefficients in VGPRs and emits every step as a copy
// plus the two-address v_fmac_f64 (twice the instructions); the same 15 FMAs in the same
// order, so the results are bit-identical.
template <bool SGPR_COEF = false> __device__ __forceinline__ double atan_poly64(double a)
{
    const double s = a * a;
    double p = -4.7926633120228785553e-5;
    p = fma(p, s, 4.800660567800612957e-4);
    if constexpr (SGPR_COEF)
    {
        asm("v_fma_f64 %0, %0, %1, %2\n\t"
            "v_fma_f64 %0, %0, %1, %3\n\t"
            "v_fma_f64 %0, %0, %1, %4\n\t"
            "v_fma_f64 %0, %0, %1, %5\n\t"
            "v_fma_f64 %0, %0, %1, %6\n\t"
            "v_fma_f64 %0, %0, %1, %7\n\t"
            "v_fma_f64 %0, %0, %1, %8\n\t"
            "v_fma_f64 %0, %0, %1, %9\n\t"
            "v_fma_f64 %0, %0, %1, %10\n\t"
            "v_fma_f64 %0, %0, %1, %11\n\t"
            "v_fma_f64 %0, %0, %1, %12\n\t"
            "v_fma_f64 %0, %0, %1, %13\n\t"
            "v_fma_f64 %0, %0, %1, %14\n\t"
            "v_fma_f64 %0, %0, %1, %15\n\t"
            "v_fma_f64 %0, %0, %1, %16"
            : "+v"(p)
            : "v"(s), "s"(-2.2725728333658927666e-3), "s"(6.8183912427924839848e-3),
              "s"(-1.4739706399322470277e-2), "s"(2.4909137877934616446e-2), "s"(-3.5142229350564300719e-2),
              "s"(4.3938991765738211806e-2), "s"(-5.1356486201903778769e-2), "s"(5.8483233451193392607e-2),
              "s"(-6.6598275562737410628e-2), "s"(7.6913027337110721971e-2), "s"(-9.090805442522420044e-2),
              "s"(1.1111104041852547122e-1), "s"(-1.4285713995249626285e-1), "s"(1.9999999993898812193e-1),
              "s"(-3.3333333333288106455e-1));
        return fma(a * s, p, a);
    }
    p = fma(p, s, -2.2725728333658927666e-3);
    p = fma(p, s, 6.8183912427924839848e-3);
    p = fma(p, s, -1.4739706399322470277e-2);
    p = fma(p, s, 2.4909137877934616446e-2);
    p = fma(p, s, -3.5142229350564300719e-2);
    p = fma(p, s, 4.3938991765738211806e-2);
    p = fma(p, s, -5.1356486201903778769e-2);
    p = fma(p, s, 5.8483233451193392607e-2);
    p = fma(p, s, -6.6598275562737410628e-2);
    p = fma(p, s, 7.6913027337110721971e-2);
    p = fma(p, s, -9.090805442522420044e-2);
    p = fma(p, s, 1.1111104041852547122e-1);
    p = fma(p, s, -1.4285713995249626285e-1);
    p = fma(p, s, 1.9999999993898812193e-1);
    p = fma(p, s, -3.3333333333288106455e-1);
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
// (the hardware reciprocal, within 2^-24 relative on gfx950 (measured), refined by one
// Newton step to ~2^-48: a within ~4e-15 relative, and atan(a) within that as well since
// a / ((1 + a^2) atan(a)) <= 1), atan(a) from the polynomial, then the octant:
// pi/2 - . when |y| > |x|, pi - . when x < 0, sign of y.
__device__ __forceinline__ double pvl_atan2(double y, double x)
{
    const double ax = fabs(x), ay = fabs(y);
    const double mn = fmin(ax, ay), mx = fmax(ax, ay);
    double r = __builtin_amdgcn_rcp(mx);
    r = fma(r, fma(-mx, r, 1.0), r);
    const double a = mn * r;
    double t = atan_poly64(a);
    t = (ay > ax) ? (1.57079632679489661923 - t) : t;
    t = (x < 0.0) ? (3.14159265358979323846 - t) : t;
    return copysign(t, y);
}

// atan2(-ay, x) for ay >= 0 (the segment's angle below: its sine is never positive), as
// pvl_atan2 without the sign transfer.
template <bool SGPR_COEF = false> __device__ __forceinline__ double pvl_atan2_neg(double ay, double x)
{
    const double ax = fabs(x);
    const double mn = fmin(ax, ay), mx = fmax(ax, ay);
    double r = __builtin_amdgcn_rcp(mx);
    r = fma(r, fma(-mx, r, 1.0), r);
    const double a = mn * r;
    double t = atan_poly64<SGPR_COEF>(a);
    t = (ay > ax) ? (1.57079632679489661923 - t) : t;
    t = (x < 0.0) ? (3.14159265358979323846 - t) : t;
    return -t;
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

// 1/sqrt(x) for x > 0 normal: the hardware estimate y (within 2^-23 relative on gfx950,
// measured) and one third-order step of 1/sqrt(x) = y (1 - e)^(-1/2) =
// y (1 + e/2 + 3 e^2/8 + 5 e^3/16 + ...), e = 1 - x y^2: the first omitted term is below
// 2^-66 relative for |e| <= 2^-22, so the result is within rounding of 1/sqrt(x).
__device__ __forceinline__ double pvl_rsqrt(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y), y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Scalar factor of K from the pair's endpoint terms: dr1 = p - r1 and q1 = |dr1|^2, dr2 and
// q2 likewise (each dot product left to right, no FMA, as the reference), e = the unit
// direction.  The cut-off is decided here.  Split from line_m so that a walk along a chain of
// segments can reuse one endpoint's terms for the next segment (bit-identical values).
template <int PREC, bool SGPR_COEF = false>
__device__ __forceinline__ double pair_m(double ex, double ey, double ez, double len, double d1x, double d1y,
                                         double d1z, double q1, double d2x, double d2y, double d2z, double q2,
                                         double tol, double tol2, bool &cut)
{
    const double t1 = ex * d1x + ey * d1y + ez * d1z;
    const double t2 = ex * d2x + ey * d2y + ez * d2z;
    const double n1 = q1 - t1 * t1;
    const double n2 = q2 - t2 * t2;
    const double h2 = (n1 + n2) / 2.0;
    double m;
    if constexpr (PREC == PVL_FP64_REF_FORM)
    {
        cut = (len < tol) || (h2 < tol2);
        const double hh = cut ? 1.0 : h2;
        const double h = sqrt(hh);
        m = (atan2(t2, h) - atan2(t1, h)) / h;
    }
    else
    {
        // Only h2 < tol^2 is tested here: a line shorter than tol has direction 0 and length 0
        // (line_geom_kernel), so t1 = t2 = 0, the angle atan2(-0, h2 > 0) is 0 and the line
        // adds exactly 0, as the reference's cut-off does.  A cut pair's h2 (possibly <= 0)
        // may give NaN or Inf below; the caller replaces the pair's value by 0.
        // t2 - t1 = e . (r1 - r2) = -len exactly in exact arithmetic; -len is used (the
        // rounded difference of t1 and t2 carries an error of eps |dr| / len relative in the
        // far field).  A line cut by length has len = 0 (line_geom_kernel), so its angle is 0.
        cut = h2 < tol2;
        if constexpr (PREC == PVL_FP64)
        {
            const double ih = pvl_rsqrt(h2);
            const double h = h2 * ih;
            const double dphi = pvl_atan2_neg<SGPR_COEF>(len * h, fma(t1, t2, h2));
            m = dphi * ih;
        }
        else
        {
            const float hf = sqrtf((float)h2);
            const float ihf = __builtin_amdgcn_rcpf(hf);
            const float num = -(float)len * hf;
            const float den = (float)fma(t1, t2, h2);
            m = (double)(pvl_atan2f(num, den) * ihf);
        }
    }
    return m;
}

// K = c m for the wake's self-induction in the accurate forms (PVL_FP64, PVL_FP32), from
// dr1 = p - r1 and the segment's unit direction e and length: c = dr1 x e (K's direction,
// |c| = h, the distance of p from the segment's line), h^2 = |c|^2, t1 = e . dr1 (FMAs),
// t2 = t1 - len; returns m = angle / h and decides the cut-off (h^2 < tol^2).  Mathematically
// the reference's h^2; numerically its error is ~eps |dr1| / h relative (the cross product's
// own, which K's direction carries in every form), where the reference's |dr|^2 - t^2 loses
// ~eps |dr|^2 / h^2, so the two differ by the reference's rounding, which is of the order of
// eps wherever |dr| / h is of the order of one (a sheet of segments of similar length, as a
// free wake is).  It needs neither |dr1|^2 nor dr2, ~18 instructions fewer per pair.  A line
// shorter than tol has direction 0 and length 0 (chain records), so c = 0, h^2 = 0 and it is
// cut, as the reference's length test cuts it.  A cut pair's m may be NaN or Inf; the caller
// replaces the pair's value by 0.
template <int PREC>
__device__ __forceinline__ double cross_m(double ex, double ey, double ez, double len, double d1x, double d1y,
                                          double d1z, double tol2, double &cx, double &cy, double &cz, bool &cut)
{
    cx = fma(d1y, ez, -(d1z * ey));
    cy = fma(d1z, ex, -(d1x * ez));
    cz = fma(d1x, ey, -(d1y * ex));
    const double h2 = fma(cx, cx, fma(cy, cy, cz * cz));
    const double t1 = fma(ex, d1x, fma(ey, d1y, ez * d1z));
    const double t2 = t1 - len;
    cut = h2 < tol2;
    if constexpr (PREC == PVL_FP64)
    {
        const double ih = pvl_rsqrt(h2);
        const double h = h2 * ih;
        return pvl_atan2_neg(len * h, fma(t1, t2, h2)) * ih;
    }
    else
    {
        const float hf = sqrtf((float)h2);
        const float ihf = __builtin_amdgcn_rcpf(hf);
        const float num = -(float)len * hf;
        const float den = (float)fma(t1, t2, h2);
        return (double)(pvl_atan2f(num, den) * ihf);
    }
}

// Scalar factor of K (angle / h; the cut-off is decided here) and dr1 = p - r1.
template <int PREC, bool SGPR_COEF = false>
__device__ __forceinline__ double line_m(const LineGeom &g, double px, double py, double pz, double tol, double tol2,
                                         double &d1x, double &d1y, double &d1z, bool &cut)
{
    d1x = px - g.r1x, d1y = py - g.r1y, d1z = pz - g.r1z;
    const double d2x = px - g.r2x, d2y = py - g.r2y, d2z = pz - g.r2z;
    const double q1 = d1x * d1x + d1y * d1y + d1z * d1z;
    const double q2 = d2x * d2x + d2y * d2y + d2z * d2z;
    return pair_m<PREC, SGPR_COEF>(g.ex, g.ey, g.ez, g.len, d1x, d1y, d1z, q1, d2x, d2y, d2z, q2, tol, tol2, cut);
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

// n . K(line, p) for the normal-reduced assembly.  The reference form dots the reference's K
// (reference rounding of every product); the others form m (n . c) with c = dr1 x e and the
// dot as FMAs, and decide the cut-off once for the scalar.
template <int PREC>
__device__ __forceinline__ double line_nk(const LineGeom &g, double px, double py, double pz, double nx, double ny,
                                          double nz, double tol, double tol2)
{
    if constexpr (PREC == PVL_FP64_REF_FORM)
    {
        double kx, ky, kz;
        line_k<PREC>(g, px, py, pz, tol, tol2, kx, ky, kz);
        return kx * nx + ky * ny + kz * nz;
    }
    else
    {
        double d1x, d1y, d1z;
        bool cut;
        const double m = line_m<PREC, true>(g, px, py, pz, tol, tol2, d1x, d1y, d1z, cut);
        const double cx = fma(d1y, g.ez, -(d1z * g.ey));
        const double cy = fma(d1z, g.ex, -(d1x * g.ez));
        const double cz = fma(d1x, g.ey, -(d1y * g.ex));
        const double nc = fma(cx, nx, fma(cy, ny, cz * nz));
        return cut ? 0.0 : m * nc;
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

