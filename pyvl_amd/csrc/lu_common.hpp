// Helpers shared by the LU translation units.
#pragma once

#include <hip/hip_runtime.h>

namespace pvl
{
// Pivot order of LAPACK's idamax: larger |a| first, the smaller row on ties (a NaN never
// wins, so a column of NaNs keeps its diagonal row).
__device__ __forceinline__ bool better(double v1, int p1, double v2, int p2)
{
    return v1 > v2 || (v1 == v2 && p1 < p2);
}
} // namespace pvl
