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
// registers, for panels short enough to fit (m <= 512 RPT): thread t owns panel rows
// t + 512 s, s < RPT.  The same arithmetic and pivot choice as LAPACK's dgetf2 (and
// panel_kernel), but a column costs two workgroup barriers instead of a grid-wide exchange:
// DPP wave argmax -> 8-entry LDS reduction -> the pivot row's owner publishes it in LDS ->
// swap, scale by the reciprocal pivot, rank-1 update.
//
// Each row is a rotating window of NBP registers so that the column loop can stay a
// runtime loop with static register indices: before step J, window slot w holds column
// J + w for w < NBP - J and L(., w - NBP + J) after that.  A step consumes slot 0 (the
// current column), shifts the trailing slots left while applying the update, and appends
// the new L value in the last slot.  Rows swap whole windows (dgetf2 swaps whole rows), a
// pivot row freezes at its step D, and the write-back puts slot w in column (w + D) mod NBP.
template <typename T, int NBP, int RPT>
__global__ void __launch_bounds__(512) panel_wg_kernel(T *__restrict__ P, int lda, int m, int jb, int row_base,
                                                       int *__restrict__ piv, int *__restrict__ status)
{
    main_stream_priority<2>();
    static_assert((NBP & (NBP - 1)) == 0, "NBP must be a power of two");
    constexpr int NT = 512;
    constexpr unsigned NONE = 0xFFFFFFFFu;
    __shared__ T prow[2][NBP];
    __shared__ T rowj[2][NBP];
    __shared__ unsigned long long red_k[2][NT / 64];
    __shared__ unsigned red_p[2][NT / 64];
    // pivots stay in LDS until the end: a global store inside the loop would make every
    // barrier wait for it (the workgroup fence of __syncthreads covers global memory)
    __shared__ int spiv[NBP];
    int first_zero = 0; // tid 0: first exactly-zero pivot (1-based column), 0 if none
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    T a[RPT][NBP];
    int done[RPT];
#pragma unroll
    for (int s = 0; s < RPT; ++s)
    {
        const int r = tid + NT * s;
        done[s] = jb;
#pragma unroll
        for (int c = 0; c < NBP; ++c)
            a[s][c] = (r < m && c < jb) ? P[(size_t)r * lda + c] : T(0);
    }
    // Wave-level argmax of column J (window slot 0) among rows >= J, left in red[J & 1].
    // For J >= 1 it runs at the end of the previous column's step, in the same basic block
    // as the tail of its rank-1 update, so the DPP chain overlaps those FMAs.
    auto candidate = [&](int J) {
        // rows ascend with s, so a tie keeps the first
        unsigned long long key = 0;
        unsigned row = NONE;
#pragma unroll
        for (int s = 0; s < RPT; ++s)
        {
            const int r = tid + NT * s;
            const double v = (double)a[s][0];
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
            red_k[J & 1][w] = wk, red_p[J & 1][w] = wr;
    };
    candidate(0);
    for (int J = 0; J < jb; ++J)
    {
        const int par = J & 1;
        __syncthreads();
        unsigned long long bk = 0;
        unsigned br = NONE;
#pragma unroll
        for (int q = 0; q < NT / 64; ++q)
        {
            const unsigned long long k = red_k[par][q];
            const unsigned r = red_p[par][q];
            if (k > bk || (k == bk && r < br))
                bk = k, br = r;
        }
        const int p = br == NONE ? J : (int)br; // every candidate NaN: keep row J in place
        // row J lives in thread J (J < NBP <= 512), slot 0; the pivot row in thread p % 512,
        // slot p / 512.  Both windows are in the same state (both rows are still active).
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
        if (tid == J)
            done[0] = J; // row J is now the pivot row: U(J, J..) and L(J, ..J-1), final
        if (tid == 0)
            spiv[J] = row_base + p;
        const T pv = prow[par][0];
        const bool zero = !(pv != T(0));
        if (pv == T(0) && tid == 0 && first_zero == 0)
            first_zero = row_base + J + 1;
        // dgetf2 skips the scaling for a zero pivot (the column below is then zero too) but
        // still applies the (null) rank-1 update; the window shift happens either way
        const T inv = zero ? T(1) : T(1) / pv;
#pragma unroll
        for (int s = 0; s < RPT; ++s)
        {
            const int r = tid + NT * s;
            if (r < m && r > J)
            {
                const T l = zero ? a[s][0] : a[s][0] * inv;
#pragma unroll
                for (int c = 0; c + 1 < NBP; ++c)
                {
                    // slot c + 1 is a trailing column while c + 1 < NBP - J, else an L entry
                    const T u = (c + 1 < NBP - J) ? prow[par][c + 1] : T(0);
                    a[s][c] = fma(-l, u, a[s][c + 1]);
                }
                a[s][NBP - 1] = l;
            }
        }
        if (J + 1 < jb) // uniform
            candidate(J + 1);
    }
    __syncthreads();
    if (tid < jb)
        piv[row_base + tid] = spiv[tid];
    if (tid == 0 && first_zero)
        atomicCAS(&status[0], 0, first_zero); // LAPACK info: the first zero pivot
#pragma unroll
    for (int s = 0; s < RPT; ++s)
    {
        const int r = tid + NT * s;
        if (r < m)
        {
            const int d = done[s];
#pragma unroll
            for (int c = 0; c < NBP; ++c)
            {
                const int col = (c + d) & (NBP - 1);
                if (col < jb)
                    P[(size_t)r * lda + col] = a[s][c];
            }
        }
    }
}

} // namespace lu_panel
} // namespace pvl
