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

// Wave-wide reductions through DPP lane moves (no LDS round trips).  Each step combines a
// lane with the lane 1, 2, 4, 8 places below it inside its row of 16 (row_shr), then rows
// 0 -> 1 and 2 -> 3 (row_bcast:15), then rows 0-1 -> 2-3 (row_bcast:31); lane 63 ends with
// the reduction of all 64 lanes and readlane broadcasts it.  Lanes with no source keep the
// identity, so the operation only needs an identity element.
template <int CTRL, int ROW_MASK> __device__ __forceinline__ unsigned dpp_u32(unsigned identity, unsigned v)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)identity, (int)v, CTRL, ROW_MASK, 0xF, false);
}

template <int CTRL, int ROW_MASK> __device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v)
{
    const unsigned lo = dpp_u32<CTRL, ROW_MASK>(0u, (unsigned)v);
    const unsigned hi = dpp_u32<CTRL, ROW_MASK>(0u, (unsigned)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ unsigned long long umax64(unsigned long long a, unsigned long long b)
{
    return a > b ? a : b;
}

// max over the wave of a 64-bit unsigned value (identity 0), in every lane.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v)
{
    v = umax64(v, dpp_u64<0x111, 0xF>(v)); // row_shr:1
    v = umax64(v, dpp_u64<0x112, 0xF>(v)); // row_shr:2
    v = umax64(v, dpp_u64<0x114, 0xF>(v)); // row_shr:4
    v = umax64(v, dpp_u64<0x118, 0xF>(v)); // row_shr:8
    v = umax64(v, dpp_u64<0x142, 0xA>(v)); // row_bcast:15 into rows 1 and 3
    v = umax64(v, dpp_u64<0x143, 0xC>(v)); // row_bcast:31 into rows 2 and 3
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)v, 63);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(v >> 32), 63);
    return ((unsigned long long)hi << 32) | lo;
}

// min over the wave of a 32-bit unsigned value (identity 0xFFFFFFFF), in every lane.
__device__ __forceinline__ unsigned wave_min_u32(unsigned v)
{
    constexpr unsigned I = 0xFFFFFFFFu;
    v = min(v, dpp_u32<0x111, 0xF>(I, v));
    v = min(v, dpp_u32<0x112, 0xF>(I, v));
    v = min(v, dpp_u32<0x114, 0xF>(I, v));
    v = min(v, dpp_u32<0x118, 0xF>(I, v));
    v = min(v, dpp_u32<0x142, 0xA>(I, v));
    v = min(v, dpp_u32<0x143, 0xC>(I, v));
    return __builtin_amdgcn_readlane(v, 63);
}

// Order-preserving key of a pivot candidate: the bits of |a| (monotonic for non-negative
// doubles); NaN maps to 0 and never wins (its row is excluded from the tie-break below).
__device__ __forceinline__ unsigned long long pivot_key(double v)
{
    const double a = fabs(v);
    return a == a ? (unsigned long long)__double_as_longlong(a) : 0ull;
}

// Wave-wide idamax over (key, row) pairs: the largest key, then the smallest row among the
// lanes holding it (rows of excluded candidates are 0xFFFFFFFF).  Returns the row (or
// 0xFFFFFFFF if no lane has a candidate) and the key through *kmax.
__device__ __forceinline__ unsigned wave_argmax(unsigned long long key, unsigned row, unsigned long long *kmax)
{
    const unsigned long long m = wave_max_u64(key);
    *kmax = m;
    return wave_min_u32(key == m ? row : 0xFFFFFFFFu);
}
} // namespace pvl
