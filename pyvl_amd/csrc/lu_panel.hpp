// Single-workgroup, register-resident LU panel kernel (see panel_wg_kernel), instantiated
// once per precision in lu_panel_f64.hip / lu_panel_f32.hip: the column steps are expanded
// at compile time, which makes each instantiation a large translation unit.
#pragma once

#include "lu_common.hpp"
#include "pvl_internal.hpp"

#include <climits>

namespace pvl
{
namespace lu_panel
{

// Panel factorisation by ONE workgroup of 512 threads with the whole m x jb panel held in
// registers (thread t owns panel rows t + 512 s, s < RPT; NBP >= jb columns each), for
// panels short enough to fit (m <= 512 RPT).  The same arithmetic and pivot choice as
// panel_kernel, but a column costs two workgroup barriers instead of a grid-wide exchange:
// DPP wave argmax -> 8-entry LDS reduction -> the pivot row's owner publishes it in LDS ->
// swap, scale by the reciprocal pivot, rank-1 update of the owned rows.  The columns are
// expanded at compile time (PanelStep<J>) so that every register index is static.
struct PanelShared
{
    unsigned long long red_k[2][8];
    unsigned red_p[2][8];
};

template <typename T, int NBP, int RPT, int J> struct PanelStep
{
    static __device__ __forceinline__ void run(T (&a)[RPT][NBP], T (*prow)[NBP], T (*rowj)[NBP], PanelShared &sh,
                                               int tid, int m, int jb, int row_base, int *piv, int *status)
    {
        constexpr int NT = 512;
        if (J < jb) // uniform; a skipped column has no barrier either
        {
            constexpr int par = J & 1;
            const int lane = tid & 63, w = tid >> 6;
            constexpr unsigned NONE = 0xFFFFFFFFu;
            // this thread's best candidate (rows ascend with s, so a tie keeps the first)
            unsigned long long key = 0;
            unsigned row = NONE;
#pragma unroll
            for (int s = 0; s < RPT; ++s)
            {
                const int r = tid + NT * s;
                const double v = (double)a[s][J];
                if (r < m && r >= J && v == v)
                {
                    const unsigned long long k = pivot_key(v);
                    if (row == NONE || k > key)
                        key = k, row = (unsigned)r;
                }
            }
            unsigned long long wk;
            const unsigned wr = wave_argmax(key, row, &wk);
            if (lane == 0)
                sh.red_k[par][w] = wk, sh.red_p[par][w] = wr;
            __syncthreads();
            unsigned long long bk = 0;
            unsigned br = NONE;
#pragma unroll
            for (int q = 0; q < NT / 64; ++q)
            {
                const unsigned long long k = sh.red_k[par][q];
                const unsigned r = sh.red_p[par][q];
                if (k > bk || (k == bk && r < br))
                    bk = k, br = r;
            }
            const int p = br == NONE ? J : (int)br; // every candidate NaN: keep row J in place
            // row J lives in thread J (J < 512), slot 0; the pivot row in thread p % 512, slot p / 512
#pragma unroll
            for (int s = 0; s < RPT; ++s)
                if (tid + NT * s == p)
                {
#pragma unroll
                    for (int c = 0; c < NBP; ++c)
                        prow[par][c] = a[s][c];
                }
            if (tid == J)
            {
#pragma unroll
                for (int c = 0; c < NBP; ++c)
                    rowj[par][c] = a[0][c];
            }
            __syncthreads();
            if (p != J)
            {
                if (tid == J)
                {
#pragma unroll
                    for (int c = 0; c < NBP; ++c)
                        a[0][c] = prow[par][c];
                }
#pragma unroll
                for (int s = 0; s < RPT; ++s)
                    if (tid + NT * s == p)
                    {
#pragma unroll
                        for (int c = 0; c < NBP; ++c)
                            a[s][c] = rowj[par][c];
                    }
            }
            if (tid == 0)
                piv[row_base + J] = row_base + p;
            const T pv = prow[par][J];
            if (pv != T(0))
            {
                const T inv = T(1) / pv;
#pragma unroll
                for (int s = 0; s < RPT; ++s)
                {
                    const int r = tid + NT * s;
                    if (r < m && r > J)
                    {
                        const T l = a[s][J] * inv;
                        a[s][J] = l;
#pragma unroll
                        for (int c = J + 1; c < NBP; ++c)
                            a[s][c] = fma(-l, prow[par][c], a[s][c]);
                    }
                }
            }
            else if (tid == 0)
            {
                atomicCAS(&status[0], 0, row_base + J + 1); // first exactly-zero pivot (LAPACK info)
            }
        }
        PanelStep<T, NBP, RPT, J + 1>::run(a, prow, rowj, sh, tid, m, jb, row_base, piv, status);
    }
};
template <typename T, int NBP, int RPT> struct PanelStep<T, NBP, RPT, NBP>
{
    static __device__ __forceinline__ void run(T (&)[RPT][NBP], T (*)[NBP], T (*)[NBP], PanelShared &, int, int, int,
                                               int, int *, int *)
    {
    }
};

template <typename T, int NBP, int RPT>
__global__ void __launch_bounds__(512) panel_wg_kernel(T *__restrict__ P, int lda, int m, int jb, int row_base,
                                                       int *__restrict__ piv, int *__restrict__ status)
{
    constexpr int NT = 512;
    __shared__ T prow[2][NBP];
    __shared__ T rowj[2][NBP];
    __shared__ PanelShared sh;
    const int tid = threadIdx.x;
    T a[RPT][NBP];
#pragma unroll
    for (int s = 0; s < RPT; ++s)
    {
        const int r = tid + NT * s;
#pragma unroll
        for (int c = 0; c < NBP; ++c)
            a[s][c] = (r < m && c < jb) ? P[(size_t)r * lda + c] : T(0);
    }
    PanelStep<T, NBP, RPT, 0>::run(a, prow, rowj, sh, tid, m, jb, row_base, piv, status);
#pragma unroll
    for (int s = 0; s < RPT; ++s)
    {
        const int r = tid + NT * s;
        if (r < m)
        {
#pragma unroll
            for (int c = 0; c < NBP; ++c)
                if (c < jb)
                    P[(size_t)r * lda + c] = a[s][c];
        }
    }
}

} // namespace lu_panel
} // namespace pvl
