// The line-vortex kernel K(line, p) of the reference (src/core/flow_solver.c:114-135):
//
//   dr1 = p - r1, dr2 = p - r2, t_i = e . dr_i, h^2 = ((|dr1|^2 - t1^2) + (|dr2|^2 - t2^2)) / 2
//   K = 0                                            if |d| < tol or h^2 < tol^2
//   K = cross(dr1, e) * (atan2(t2, h) - atan2(t1, h)) / h   otherwise
//
// |K| is the angle the segment subtends from p (no 1/(4 pi) factor).  The file must be
// compiled with -ffp-contract=off: the FP64 form below then rounds exactly as the
// reference's GCC x86-64 build (dot products left to right, no FMA), except for the
// last-bit differences between this library's atan2 and glibc's.
#pragma once

#include "pvl_internal.hpp"

// PVL_FP64: reference operation order, two FP64 atan2 calls.
// PVL_FP32: the cut-off test, the dot products, h^2 and cross(dr1, e) stay in FP64; the
//   angle is one FP32 atan2 of the equivalent single-angle form
//     atan2(t2, h) - atan2(t1, h) = atan2((t2 - t1) h, h^2 + t1 t2)
//   (both sides lie in (-pi, pi): sin and cos of the difference are (t2 - t1) h / (r1 r2)
//   and (h^2 + t1 t2) / (r1 r2) with r_i = sqrt(h^2 + t_i^2) > 0), which has no
//   cancellation in the far field, so the FP32 angle keeps ~1e-7 relative accuracy.
template <int PREC>
__device__ __forceinline__ void line_k(const LineGeom &g, double px, double py, double pz, double tol, double tol2,
                                       double &kx, double &ky, double &kz)
{
    const double d1x = px - g.r1x, d1y = py - g.r1y, d1z = pz - g.r1z;
    const double d2x = px - g.r2x, d2y = py - g.r2y, d2z = pz - g.r2z;
    const double t1 = g.ex * d1x + g.ey * d1y + g.ez * d1z;
    const double t2 = g.ex * d2x + g.ey * d2y + g.ez * d2z;
    const double n1 = (d1x * d1x + d1y * d1y + d1z * d1z) - t1 * t1;
    const double n2 = (d2x * d2x + d2y * d2y + d2z * d2z) - t2 * t2;
    const double h2 = (n1 + n2) / 2.0;
    const bool cut = (g.len < tol) || (h2 < tol2);
    const double hh = cut ? 1.0 : h2;
    double m;
    if constexpr (PREC == PVL_FP64)
    {
        const double h = sqrt(hh);
        m = (atan2(t2, h) - atan2(t1, h)) / h;
    }
    else
    {
        const float hf = sqrtf((float)hh);
        const float num = (float)(t2 - t1) * hf;
        const float den = (float)(hh + t1 * t2);
        m = (double)(atan2f(num, den) / hf);
    }
    const double cx = d1y * g.ez - d1z * g.ey;
    const double cy = d1z * g.ex - d1x * g.ez;
    const double cz = d1x * g.ey - d1y * g.ex;
    kx = cut ? 0.0 : cx * m;
    ky = cut ? 0.0 : cy * m;
    kz = cut ? 0.0 : cz * m;
}
