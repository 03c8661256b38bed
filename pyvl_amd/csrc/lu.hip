// Dense LU with partial pivoting on gfx950 (FP64 and FP32), triangular solves, FP64
// residuals and mixed-precision iterative refinement.
//
// The reference solves with scipy.linalg.lu_factor / lu_solve (python/pyvl/solver.py:223-227),
// i.e. LAPACK dgetrf/dgetrs: P A = L U with the first largest |a| of each column as pivot.
// Here A is row-major n x n (lda = n): the same factorisation as LAPACK applied to the
// same matrix, with piv[j] = row swapped with row j at step j (0-based, as scipy returns).
//
// Right-looking blocked algorithm, panel width NB = 64:
//   panel_kernel   grid-wide; each workgroup holds up to R panel rows in LDS; per
//                  column every group publishes its pivot candidate's row and then a tagged
//                  record of the candidate, and reads every group's record (no barrier).
//   laswp_kernel   applies the panel's interchanges to the other columns.
//   trsm_kernel    U12 = L11^-1 A12 (one lane per column in registers, L11 in LDS).
//   gemm_kernel    A22 -= L21 U12 on MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32)
//                  for n <= 2048 and the short updates of larger systems; the trailing
//                  update after each outer block of 256 columns, and the updates inside an
//                  outer block with a side of 2048 or more, are plain GEMMs and go to rocBLAS
//                  (dgemm / sgemm, no atomics), which runs m x m x 256 ~1.4x faster.
// Solves: one workgroup for systems whose right-hand sides fit its LDS (trsv_wg_kernel),
// else one launch per triangle with a workgroup per block row that waits on the flags of
// the blocks before it (trsv_flow_kernel).
#include "pvl_internal.hpp"
#include "lu_common.hpp"

#include <rocblas/rocblas.h>

#include <cfloat>

#include <climits>
#include <cmath>
#include <type_traits>
#include <utility>
#include <vector>

using namespace pvl;

namespace
{

constexpr int NB = 64;        // panel / block width
constexpr int kMaxGroups = 256;
constexpr int kMaxRhs = 32;   // right-hand sides per solve launch
constexpr unsigned kSpinLimit = 20000000u; // bounded waits: ~1 s, then the call fails

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma;
// Accumulator layouts (measured on gfx950): element v of lane l holds
//   f64 16x16x4: D[4*v + l/16][l%16]      f32 16x16x4: D[4*(l/16) + v][l%16]
template <> struct Mfma<double>
{
    using acc_t = double4_t;
    static __device__ __forceinline__ int out_row(int lk, int v) { return 4 * v + lk; }
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<float>
{
    using acc_t = float4_t;
    static __device__ __forceinline__ int out_row(int lk, int v) { return 4 * lk + v; }
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

template <typename T> struct Bits;
template <> struct Bits<double>
{
    using U = unsigned long long;
};
template <> struct Bits<float>
{
    using U = unsigned int;
};

// Cache-bypassing loads/stores for the grid-wide exchange buffers.
template <typename T> __device__ __forceinline__ T xload(const T *p)
{
    using U = typename Bits<T>::U;
    U u = __hip_atomic_load((const U *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(T, u);
}
template <typename T> __device__ __forceinline__ void xstore(T *p, T v)
{
    using U = typename Bits<T>::U;
    __hip_atomic_store((U *)p, __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int xload_i(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xstore_i(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long xload_u64(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xstore_u64(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned xload_u32(const unsigned *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void xstore_u32(unsigned *p, unsigned v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded waits: a wait that has not ended after kWaitTicks of the 100 MHz real-time
// counter (or once another wave has flagged a timeout) gives up and flags *err.
constexpr unsigned long long kWaitTicks = 200000000ull; // 2 s
__device__ __forceinline__ bool wait_expired(unsigned &spins, unsigned long long t0, const int *err)
{
    if ((++spins & 255u) != 0)
        return false;
    return wall_clock64() - t0 > kWaitTicks || xload_i(err) != 0;
}

constexpr int kRecBytes = 256;

// Exchange area for the panel kernel, double buffered by the parity of the exchange round
// (one round per column).  There is no grid barrier: a group's record of a round is the
// signal that its published rows of that round are complete, and every word of a record
// carries a tag bit (bit 1 of the round number), so that a reader tells the record of this
// round from the one two rounds earlier in the same buffer.  A group can run at most one
// round ahead of the slowest (it needs every group's record of the round before), so two
// buffers and one tag bit suffice.
// Each record sits in a cache line of its own: every group polls every record, and records
// packed into a few lines queue those loads (and the stores behind them) on one memory channel.
struct alignas(kRecBytes) PanelSlot
{
    unsigned long long key; // bit 63: tag; bits 0-62: the candidate's |a| (double bits)
    unsigned pos;           // bit 31: tag; bits 0-30: its panel position (INT_MAX: none)
};
struct PanelRec
{
    PanelSlot slot[2][kMaxGroups];
};
template <typename T> struct PanelX
{
    PanelRec rec;
    T crow[2][kMaxGroups][NB];
    T rowj[2][NB];
};

// ---------------------------------------------------------------------------------------
// The panel's columns [c0, jb) of local rows [rb, nr) take a sub-block's update,
//   x[r][c] -= sum_i L[r][j0 + i] * U[i][c]   (i = 0 .. kb-1 ascending, zero pivots skipped),
// each element by the same FMAs in the same order as the column-by-column sweep.  A thread
// takes TR rows x 8 columns: per FMA it reads a quarter (TR = 8) of the operands that one
// row per thread with every U value broadcast from LDS reads, and the LDS, not the FMA
// pipes, bounds this step.  Rows of a tile are strided by the number of row tiles, so the
// lanes of a wave read consecutive rows (row stride 2 banks: conflict-free).  Loads clamp
// their indices and stores outside the block go to row `sink` (one row past the group's
// rows) or to the padding column NBW, so the code has no per-element branches; the
// operands of step i + 1 are fetched while step i's FMAs run.
template <typename T, int TR, int KB, int LDP, int NBW>
__device__ __forceinline__ void subblock_update(T *__restrict__ L, const T (*__restrict__ U)[NBW], int rb, int nr,
                                                int sink, int j0, int kb, int c0, int jb, unsigned zm)
{
    const int nrows = nr - rb, ncols = jb - c0;
    const int nrt = (nrows + TR - 1) / TR, nct = (ncols + 7) / 8;
    const int tid = threadIdx.x;
    if (tid >= nrt * nct)
        return;
    const int rt = tid % nrt, ct = tid / nrt;
    const int cb = c0 + 8 * ct;
    int rl[TR], rs[TR], cl[8], cs[8];
#pragma unroll
    for (int q = 0; q < TR; ++q)
    {
        const int r = rb + rt + nrt * q;
        rl[q] = min(r, nr - 1) * LDP;
        rs[q] = (r < nr ? r : sink) * LDP;
    }
#pragma unroll
    for (int cc = 0; cc < 8; ++cc)
    {
        cl[cc] = min(cb + cc, jb - 1);
        cs[cc] = cb + cc < jb ? cb + cc : NBW;
    }
    T x[TR][8];
#pragma unroll
    for (int q = 0; q < TR; ++q)
#pragma unroll
        for (int cc = 0; cc < 8; ++cc)
            x[q][cc] = L[rl[q] + cl[cc]];
    T mu[TR], u[8];
#pragma unroll
    for (int q = 0; q < TR; ++q)
        mu[q] = L[rl[q] + j0];
#pragma unroll
    for (int cc = 0; cc < 8; ++cc)
        u[cc] = U[0][cl[cc]];
#pragma unroll
    for (int i = 0; i < KB; ++i)
    {
        T mn[TR], un[8];
        if (i + 1 < KB)
        {
            const int in = min(i + 1, kb - 1);
#pragma unroll
            for (int q = 0; q < TR; ++q)
                mn[q] = L[rl[q] + j0 + in];
#pragma unroll
            for (int cc = 0; cc < 8; ++cc)
                un[cc] = U[in][cl[cc]];
        }
        if (i < kb && !((zm >> i) & 1u))
        {
#pragma unroll
            for (int q = 0; q < TR; ++q)
#pragma unroll
                for (int cc = 0; cc < 8; ++cc)
                    x[q][cc] = fma(-mu[q], u[cc], x[q][cc]);
        }
        if (i + 1 < KB)
        {
#pragma unroll
            for (int q = 0; q < TR; ++q)
                mu[q] = mn[q];
#pragma unroll
            for (int cc = 0; cc < 8; ++cc)
                u[cc] = un[cc];
        }
    }
#pragma unroll
    for (int q = 0; q < TR; ++q)
#pragma unroll
        for (int cc = 0; cc < 8; ++cc)
            L[rs[q] + cs[cc]] = x[q][cc];
}

// ---------------------------------------------------------------------------------------
// Panel factorisation of P (m x jb, row stride lda), rows k.. of the matrix.  Group g owns
// panel positions g, g + G, g + 2G, ... (at most R of them).  Per column j: local argmax
// over positions >= j; publish the candidate's row and the current row j, then the
// candidate's record (|value|, position); read every group's record; every group picks the
// same winner (largest |a|, first position on ties: LAPACK idamax), swaps, scales by the
// reciprocal pivot and applies the rank-1 update to its rows.
template <typename T>
__global__ void __launch_bounds__(256) panel_kernel(T *__restrict__ P, int lda, int m, int jb, int R, int row_base,
                                                    int *__restrict__ piv, PanelX<T> *__restrict__ X, unsigned base,
                                                    int *__restrict__ status)
{
    main_stream_priority<2>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int LDP = NB + 1;
    T *L = (T *)smem_raw;
    __shared__ T prow[NB], srowj[NB], s_invj;
    __shared__ double red_v[4];
    __shared__ int red_p[4];
    __shared__ int s_pos, s_grp;
    // pivots collected in LDS and stored once at the end: a global store before a barrier
    // makes the barrier wait for it
    __shared__ int spiv[NB];
    int first_zero = 0;

    const int tid = threadIdx.x, g = blockIdx.x, G = gridDim.x;
    // rows dealt cyclically: group g holds panel positions g, g + G, g + 2G, ... (at most R),
    // so row j, its owner's extra publishing and the sub-block's pivot rows rotate over the
    // groups instead of all falling to the first
    const int nr = g < m ? (m - 1 - g) / G + 1 : 0;
    for (int idx = tid; idx < nr * jb; idx += 256)
    {
        const int r = idx / jb, c = idx - r * jb;
        L[r * LDP + c] = P[(size_t)(g + r * G) * lda + c];
    }
    __syncthreads();

    // Columns in sub-blocks of KB: a sub-block's columns live in registers (one row per thread)
    // and are updated column by column; the panel's columns right of the sub-block are
    // updated once per sub-block from the sub-block's pivot rows (U12 of the sub-block,
    // formed by every group from the published pivot rows) -- per element the same FMAs in the
    // same order as the column-by-column sweep, so factors and pivots are unchanged, with an
    // eighth of the LDS traffic.  Rows published for the exchange carry the sub-block's
    // columns from registers and the columns right of it as they stand (not yet updated by the
    // sub-block): every consumer (the interchange, U12) treats them the same way.
    constexpr int KB = 8;
    __shared__ T ucur[KB][NB]; // the sub-block's pivot rows as published, then U12 in place
    __shared__ T wsub[4][KB], rowj_sub[KB]; // each wave's candidate's sub-block columns; row j's
    const int r = tid; // this thread's row (R <= 256 rows per group)
    const bool has_row = r < nr;
    const int pos = g + r * G;
    for (int j0 = 0; j0 < jb; j0 += KB)
    {
        const int kb = min(KB, jb - j0);
        T v[KB];
#pragma unroll
        for (int t = 0; t < KB; ++t)
            v[t] = (has_row && t < kb) ? L[r * LDP + j0 + t] : T(0);
        unsigned zm = 0; // bit t: column j0 + t had an exactly-zero pivot (no update applied)
#pragma unroll
        for (int t = 0; t < KB; ++t)
        {
            if (t >= kb)
                break;
            const int j = j0 + t;
            const unsigned rnd = base + (unsigned)j; // exchange round
            const int par = (int)(rnd & 1u);
            const unsigned tagbit = (rnd >> 1) & 1u;
            const unsigned long long ktag = (unsigned long long)tagbit << 63;
            const unsigned ptag = tagbit << 31;
            // 1. local candidate: per lane, then over the wave (DPP moves), then over the waves;
            // each wave's winner stages its sub-block columns beside the wave's record, and the
            // owner of row j its own, before the one barrier
            const int oj = j % G; // the group owning row j
            const bool own_j = g == oj;
            {
                unsigned long long key = 0;
                unsigned row = 0xFFFFFFFFu;
                if (has_row && pos >= j)
                {
                    const double vv = fabs((double)v[t]);
                    if (vv == vv)
                        key = pivot_key(vv), row = (unsigned)pos;
                }
                unsigned long long km;
                const unsigned wrow = wave_argmax(key, row, &km);
                if ((tid & 63) == 0)
                {
                    const bool none = wrow == 0xFFFFFFFFu;
                    red_v[tid >> 6] = none ? -1.0 : __longlong_as_double((long long)km);
                    red_p[tid >> 6] = none ? INT_MAX : (int)wrow;
                }
                if (has_row && (unsigned)pos == wrow)
                {
#pragma unroll
                    for (int i = 0; i < KB; ++i)
                        wsub[tid >> 6][i] = v[i];
                }
            }
            if (has_row && pos == j)
            {
#pragma unroll
                for (int i = 0; i < KB; ++i)
                    rowj_sub[i] = v[i];
            }
            __syncthreads();
            double bv = red_v[0];
            int bp = red_p[0], bw = 0;
            for (int w = 1; w < 4; ++w)
                if (better(red_v[w], red_p[w], bv, bp))
                    bv = red_v[w], bp = red_p[w], bw = w;
            // 2. publish the candidate's row and row j (if owned); once those stores are
            // complete, the tagged record that signals them
            // (a candidate that is row j itself is not published twice: readers take row j)
            if (bp != INT_MAX && bp != j)
                for (int c = tid; c < jb; c += 256)
                    xstore(&X->crow[par][g][c],
                           (c >= j0 && c < j0 + kb) ? wsub[bw][c - j0] : L[(bp / G) * LDP + c]);
            if (own_j)
                for (int c = tid; c < jb; c += 256)
                    xstore(&X->rowj[par][c], (c >= j0 && c < j0 + kb) ? rowj_sub[c - j0] : L[(j / G) * LDP + c]);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            if (tid == 0)
            {
                xstore_u64(&X->rec.slot[par][g].key,
                           (bp == INT_MAX ? 0ull : (unsigned long long)__double_as_longlong(bv)) | ktag);
                xstore_u32(&X->rec.slot[par][g].pos, (unsigned)bp | ptag);
            }
            // 3. every group's record of this round (wave 0, a group per lane) and the winner,
            // identical in every group; wave 1 fetches row j as soon as its owner's record is
            // in -- it is the pivot row whenever no interchange is due (the usual case)
            if (tid < 64)
            {
                unsigned long long key = 0;
                unsigned row = 0xFFFFFFFFu;
                int bq = 0;
                for (int q = tid; q < G; q += 64)
                {
                    unsigned long long kw;
                    unsigned pw;
                    unsigned spins = 0;
                    const unsigned long long t0 = wall_clock64();
                    for (;;)
                    {
                        kw = xload_u64(&X->rec.slot[par][q].key);
                        pw = xload_u32(&X->rec.slot[par][q].pos);
                        if ((kw >> 63) == (ktag >> 63) && (pw >> 31) == (ptag >> 31))
                            break;
                        if (wait_expired(spins, t0, &status[1]))
                        {
                            xstore_i(&status[1], 1);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                    const unsigned long long kk = kw & ~(1ull << 63);
                    const unsigned qp = pw & 0x7FFFFFFFu;
                    if (qp != (unsigned)INT_MAX && (kk > key || (kk == key && qp < row)))
                        key = kk, row = qp, bq = q;
                }
                unsigned long long km;
                const unsigned wrow = wave_argmax(key, row, &km);
                const unsigned long long who = __ballot(wrow != 0xFFFFFFFFu && row == wrow);
                const int wl = who ? (int)__builtin_ctzll(who) : 0;
                const int wgq = __builtin_amdgcn_readlane(bq, wl);
                if (tid == 0)
                    s_pos = wrow == 0xFFFFFFFFu ? INT_MAX : (int)wrow, s_grp = wgq;
            }
            else if (tid < 128)
            {
                unsigned pw;
                unsigned spins = 0;
                const unsigned long long t0 = wall_clock64();
                for (;;)
                {
                    pw = xload_u32(&X->rec.slot[par][oj].pos);
                    if ((pw >> 31) == (ptag >> 31))
                        break;
                    if (wait_expired(spins, t0, &status[1]))
                    {
                        xstore_i(&status[1], 1);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (tid - 64 < jb)
                {
                    const T rj = xload(&X->rowj[par][tid - 64]);
                    srowj[tid - 64] = rj;
                    // the reciprocal of row j's diagonal, ready for the usual case that row j
                    // is the pivot row (the division runs while wave 0 still polls)
                    if (tid - 64 == j)
                        s_invj = rj != T(0) ? T(1) / rj : T(0);
                }
            }
            __syncthreads();
            int p = s_pos;
            const int wg = s_grp;
            const bool none = (p == INT_MAX); // every candidate NaN: keep row j in place
            if (none)
                p = j;
            // the pivot row: row j itself when p == j (its owner does not publish it as a
            // candidate), else the winner's published row
            const T *prow_p = srowj;
            if (p != j)
            {
                for (int c = tid; c < jb; c += 256)
                    prow[c] = xload(&X->crow[par][wg][c]);
                __syncthreads();
                prow_p = prow;
            }
            // 5. interchange: LDS rows outside the sub-block, the owners' registers inside
            const bool own_p = p % G == g;
            if (p != j)
            {
                if (own_j)
                    for (int c = tid; c < jb; c += 256)
                        if (c < j0 || c >= j0 + kb)
                            L[(j / G) * LDP + c] = prow_p[c];
                if (own_p)
                    for (int c = tid; c < jb; c += 256)
                        if (c < j0 || c >= j0 + kb)
                            L[(p / G) * LDP + c] = srowj[c];
                if (has_row && pos == j)
                {
#pragma unroll
                    for (int i = 0; i < KB; ++i)
                        v[i] = (i < kb) ? prow_p[j0 + i] : v[i];
                }
                else if (has_row && pos == p)
                {
#pragma unroll
                    for (int i = 0; i < KB; ++i)
                        v[i] = (i < kb) ? srowj[j0 + i] : v[i];
                }
            }
            for (int c = tid; c < jb; c += 256)
                ucur[t][c] = prow_p[c];
            if (tid == 0)
                spiv[j] = row_base + p; // global store deferred: see the end of the kernel
            // 6. scale by the reciprocal pivot and update the sub-block's columns (registers)
            const T pv = prow_p[j];
            if (pv != T(0))
            {
                const T inv = p == j ? s_invj : T(1) / pv;
                if (has_row && pos > j)
                {
                    const T l = v[t] * inv;
                    v[t] = l;
#pragma unroll
                    for (int t2 = t + 1; t2 < KB; ++t2)
                        if (t2 < kb)
                            v[t2] = fma(-l, prow_p[j0 + t2], v[t2]);
                }
            }
            else
            {
                zm |= 1u << t;
                if (tid == 0 && first_zero == 0)
                    first_zero = row_base + j + 1; // first exactly-zero pivot (1-based global column)
            }
        }
        // 7. the sub-block's columns back to LDS; U12 of the sub-block for the columns right
        // of it, pivot row by pivot row in elimination order
        if (has_row)
        {
#pragma unroll
            for (int t = 0; t < KB; ++t)
                if (t < kb)
                    L[r * LDP + j0 + t] = v[t];
        }
        __syncthreads();
        for (int c = j0 + kb + tid; c < jb; c += 256)
        {
            T u[KB];
#pragma unroll
            for (int k = 0; k < KB; ++k)
            {
                if (k < kb)
                {
                    T x = ucur[k][c];
#pragma unroll
                    for (int i = 0; i < KB; ++i)
                        if (i < k && !((zm >> i) & 1u))
                            x = fma(-ucur[k][j0 + i], u[i], x);
                    u[k] = x;
                    ucur[k][c] = x;
                }
            }
        }
        __syncthreads();
        // 8. the columns right of the sub-block: rows below its pivots take its kb updates,
        // its pivot rows their U12 rows
        if (j0 + kb < jb)
        {
            const int c0 = j0 + kb, ncols = jb - c0;
            // local rows: [rp, rb) hold the sub-block's pivot positions, [rb, nr) the rows below
            const int rp = j0 - g > 0 ? (j0 - g + G - 1) / G : 0;
            const int rb = c0 - g > 0 ? (c0 - g + G - 1) / G : 0;
            const int nct = (ncols + 7) / 8;
            if (rb < nr)
            {
                if (nct >= 5)
                    subblock_update<T, 8, KB, LDP, NB>(L, ucur, rb, nr, R, j0, kb, c0, jb, zm);
                else if (nct >= 3)
                    subblock_update<T, 4, KB, LDP, NB>(L, ucur, rb, nr, R, j0, kb, c0, jb, zm);
                else if (nct == 2)
                    subblock_update<T, 2, KB, LDP, NB>(L, ucur, rb, nr, R, j0, kb, c0, jb, zm);
                else
                    subblock_update<T, 1, KB, LDP, NB>(L, ucur, rb, nr, R, j0, kb, c0, jb, zm);
            }
            for (int idx = tid; idx < (min(rb, nr) - rp) * ncols; idx += 256)
            {
                const int r = rp + idx / ncols, c = c0 + idx % ncols;
                L[r * LDP + c] = ucur[g + r * G - j0][c];
            }
        }
        __syncthreads();
    }
    if (g == 0)
    {
        for (int j = tid; j < jb; j += 256)
            piv[row_base + j] = spiv[j];
        if (tid == 0 && first_zero)
            atomicCAS(&status[0], 0, first_zero); // LAPACK info
    }
    for (int idx = tid; idx < nr * jb; idx += 256)
    {
        const int r = idx / jb, c = idx - r * jb;
        P[(size_t)(g + r * G) * lda + c] = L[r * LDP + c];
    }
}

__device__ __forceinline__ double bcast_lane(double v, int k)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, k);
    const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), k);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float bcast_lane(float v, int k)
{
    return __uint_as_float(__builtin_amdgcn_readlane(__float_as_uint(v), k));
}

// Value of lane 4 (lane / 4) + P in every lane of each quad (DPP quad_perm).
template <int P> __device__ __forceinline__ double quad_bcast(double v)
{
    constexpr int ctrl = P | (P << 2) | (P << 4) | (P << 6);
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, ctrl, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), ctrl, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int P> __device__ __forceinline__ float quad_bcast(float v)
{
    constexpr int ctrl = P | (P << 2) | (P << 4) | (P << 6);
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, false));
}

// Panel step WITHOUT interchanges (pivoting modes "none" and "auto"), fused: the LU of the
// 64 x 64 diagonal block, L21 = A21 U11^-1 and U12 = L11^-1 A12 in one launch.  A workgroup
// has 6 waves, each holding 16 rows or columns with 4 lanes per row / column (lane p of a quad
// holds entries 4k + p, k = 0..15, in registers):
//   waves 0-3  its own copy of the diagonal block (wave w: rows 16w .. 16w + 15),
//   wave 4     panel rows jb + 16g .. jb + 16g + 15 below the diagonal block (L21),
//   wave 5     columns jb + 16g .. jb + 16g + 15 of the panel's row block right of it (U12).
// Step j: the quad owning diagonal row j has published it and 1/u_jj in LDS (as soon as step
// j - 1 finalised it); after one barrier every later diagonal row and every row below forms
// its multiplier (owner lane of column j), broadcasts it over the quad (DPP) and subtracts it
// times row j, and the diagonal rows publish their multipliers l_ij, which the U12 columns
// apply one step later (x_i -= l_i,j-1 x_j-1).  The diagonal waves carry the step-to-step
// chain and run at the highest priority; the other two waves fill in behind them.  Per element the
// operations and their order are those of the right-looking column sweep (dgetf2 without
// interchanges, dtrsm).  The copies of the diagonal block agree (same arithmetic
// everywhere).  Every workgroup reads A11 at its start, so L11 \ U11 is stored by the LAST
// workgroup to finish (status[4] counts finished workgroups and is reset to 0 by that
// workgroup; no workgroup waits for another); workgroup 0 stores the identity interchanges.
// Only the last panel can be narrower than 64 (jb < 64 implies m == jb: no rows below, no
// columns right): it is padded with identity rows and columns.
// status[2]: smallest zero-pivot column (atomicMin); status[3] |= 1 when partial pivoting
// would interchange: a sub-diagonal |a| above |pivot| at its column's elimination, or a zero
// or NaN pivot.
constexpr int kNopivRows = 16;   // rows below (and columns right) per workgroup
constexpr int kNopivThreads = 384;

// Registers and LDS rings of panel_nopiv_kernel (one instance per lane).
template <typename T> struct NopivLane
{
    static constexpr int KQ = NB / 4;
    T v[KQ];       // diagonal row, row below, or column right (by role)
    T (*Urow)[NB]; // row j of U11, column 4k + p at p * KQ + k
    T (*Lcol)[NB]; // column j of L11, row 4k + p at p * KQ + k
    T *Uinv;
    int role;      // 0 diagonal block, 1 row below, 2 column right
    int p, rd, jb, row_base, first_zero;
    bool live, due;
};

// Publish diagonal row J (final once step J - 1 is done) and 1/u_JJ into ring slot J & 1.
template <int J, typename T> __device__ __forceinline__ void nopiv_publish(NopivLane<T> &L)
{
    constexpr int KQ = NopivLane<T>::KQ;
    constexpr int kj = J >> 2, pj = J & 3, s = J & 1;
    if (L.rd == J)
    {
#pragma unroll
        for (int k = 0; k < KQ; ++k)
            L.Urow[s][L.p * KQ + k] = L.v[k];
        if (L.p == pj)
            L.Uinv[s] = T(1) / L.v[kj];
    }
}

// Step J of panel_nopiv_kernel (J a compile-time constant: every register index is static).
template <int J, typename T> __device__ __forceinline__ void nopiv_step(NopivLane<T> &L)
{
    constexpr int KQ = NopivLane<T>::KQ;
    constexpr int kj = J >> 2, pj = J & 3, s = J & 1;
    const int p = L.p;
    __syncthreads();
    if (L.role != 2)
    {
        const T ujj = L.Urow[s][pj * KQ + kj];
        const T inv = L.Uinv[s];
        const double pa = fabs((double)ujj);
        T u[KQ];
#pragma unroll
        for (int k = kj; k < KQ; ++k)
            u[k] = L.Urow[s][p * KQ + k];
        if (L.role == 0)
        {
            if (J < L.jb && ujj == T(0) && L.first_zero == 0)
                L.first_zero = L.row_base + J + 1;
            if (L.rd > J)
            {
                const T own = L.v[kj];
                T l = own * inv;
                if (p == pj)
                {
                    L.due |= J < L.jb && (!(fabs((double)own) <= pa) || !(pa > 0.0));
                    L.v[kj] = l;
                    L.Lcol[s][(L.rd & 3) * KQ + (L.rd >> 2)] = l;
                }
                l = quad_bcast<pj>(l);
#pragma unroll
                for (int k = kj; k < KQ; ++k)
                {
                    const T v = fma(-l, u[k], L.v[k]);
                    L.v[k] = (k > kj || p > pj) ? v : L.v[k];
                }
            }
            if constexpr (J + 1 < NB)
                nopiv_publish<J + 1>(L);
        }
        else
        {
            const T own = L.v[kj];
            T l = own * inv;
            if (p == pj)
            {
                L.due |= L.live && (!(fabs((double)own) <= pa) || !(pa > 0.0));
                L.v[kj] = l;
            }
            l = quad_bcast<pj>(l);
#pragma unroll
            for (int k = kj; k < KQ; ++k)
            {
                const T v = fma(-l, u[k], L.v[k]);
                L.v[k] = (k > kj || p > pj) ? v : L.v[k];
            }
        }
    }
    else if constexpr (J >= 1)
    {
        // U12 columns: apply column J - 1 of L11 (published last step)
        constexpr int t = J - 1, kt = t >> 2, pt = t & 3;
        const T yt = quad_bcast<pt>(L.v[kt]);
#pragma unroll
        for (int k = kt; k < KQ; ++k)
        {
            const T v = fma(-L.Lcol[t & 1][p * KQ + k], yt, L.v[k]);
            L.v[k] = (k > kt || p > pt) ? v : L.v[k];
        }
    }
    // Keep this step's updates here: the rows below and U12 columns are otherwise only read
    // by the final stores, and the compiler would sink their FMAs past the barriers.
#pragma unroll
    for (int k = 0; k < KQ; ++k)
        asm volatile("" : "+v"(L.v[k]));
}

template <typename T, int... J> __device__ __forceinline__ void nopiv_steps(NopivLane<T> &L, std::integer_sequence<int, J...>)
{
    (nopiv_step<J>(L), ...);
}

// At most 96 VGPRs (5 waves per SIMD): two of its waves then fit on a SIMD beside four waves
// of the side stream's point-velocity kernel.
template <typename T>
__global__ void __launch_bounds__(kNopivThreads) __attribute__((amdgpu_waves_per_eu(5))) panel_nopiv_kernel(T *__restrict__ P, int lda, int m, int jb,
                                                                    int row_base, int *__restrict__ piv,
                                                                    int *__restrict__ status)
{
    constexpr int KQ = NB / 4;
    __shared__ T Urow[2][NB];
    __shared__ T Lcol[2][NB];
    __shared__ T Uinv[2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int g = blockIdx.x;
    NopivLane<T> L;
    L.Urow = Urow, L.Lcol = Lcol, L.Uinv = Uinv;
    L.role = w < 4 ? 0 : w - 3;
    L.p = lane & 3;
    L.rd = L.role == 0 ? 16 * w + (lane >> 2) : -1; // diagonal-block row of this quad
    L.jb = jb, L.row_base = row_base, L.first_zero = 0, L.due = false;
    const int rc = jb + kNopivRows * g + (lane >> 2); // panel row below / column right
    L.live = L.role != 0 && rc < m;                   // (m - jb rows below, m - jb columns right)
    const int p = L.p, rd = L.rd;
    if (L.role == 0)
        main_stream_priority<3>();
    else
        main_stream_priority<2>();
#pragma unroll
    for (int k = 0; k < KQ; ++k)
    {
        const int c = 4 * k + p;
        if (L.role == 0)
            L.v[k] = (rd < jb && c < jb) ? P[(size_t)rd * lda + c] : (rd == c ? T(1) : T(0));
        else if (L.role == 1)
            L.v[k] = L.live ? P[(size_t)rc * lda + c] : T(0);
        else
            L.v[k] = L.live ? P[(size_t)c * lda + rc] : T(0);
    }
    nopiv_publish<0>(L);
    nopiv_steps(L, std::make_integer_sequence<int, NB>{});
    if (L.role != 0 && L.live)
    {
#pragma unroll
        for (int k = 0; k < KQ; ++k)
        {
            const int c = 4 * k + p;
            if (L.role == 1)
                P[(size_t)rc * lda + c] = L.v[k];
            else
                P[(size_t)c * lda + rc] = L.v[k];
        }
    }
    if (g == 0 && L.role == 0)
    {
        if (rd < jb && p == 0)
            piv[row_base + rd] = row_base + rd;
        if (threadIdx.x == 0 && L.first_zero)
            atomicMin(&status[2], L.first_zero);
    }
    if (__ballot(L.due) && lane == 0)
        atomicOr(&status[3], 1);
    // This workgroup's read of A11 is complete (its values fed every step).  The workgroup
    // that finishes last stores L11 \ U11 over A11 and resets the counter.
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0)
        last = atomicAdd(&status[4], 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (last && L.role == 0)
    {
#pragma unroll
        for (int k = 0; k < KQ; ++k)
        {
            const int c = 4 * k + p;
            if (rd < jb && c < jb)
                P[(size_t)rd * lda + c] = L.v[k];
        }
        if (threadIdx.x == 0)
            atomicExch(&status[4], 0);
    }
}

// Row interchanges of one panel applied to every column outside it.
template <typename T>
__global__ void __launch_bounds__(256) laswp_kernel(T *__restrict__ A, int lda, int n, int k, int jb,
                                                   const int *__restrict__ piv)
{
    main_stream_priority();
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n || (c >= k && c < k + jb))
        return;
    for (int i = 0; i < jb; ++i)
    {
        const int p = piv[k + i];
        if (p != k + i)
        {
            const T a = A[(size_t)(k + i) * lda + c];
            A[(size_t)(k + i) * lda + c] = A[(size_t)p * lda + c];
            A[(size_t)p * lda + c] = a;
        }
    }
}

// U12 = L11^-1 A12 for columns [c0, c1) right of the panel.  Each wave covers 16 columns with 4
// lanes per column: lane p of a quad holds rows p, p + 4, .. of its column in registers.  Step
// t broadcasts the final x_t from its owner lane over the quad (DPP) and every lane subtracts
// l_it x_t from its rows i > t, with the l_it each lane needs contiguous in LDS
// (Lq[t][p * KP + k] = l_(4k+p),t).  The four segments of a row start KP = KQ + 2 elements
// apart, so the four addresses a step's read takes fall on distinct LDS banks (KQ apart they
// all fell on one), and the row stride is odd, so the transposed fill spreads its stores over
// the banks.  Per element the operations are those of dtrsm's column sweep.  No barrier after
// the load.  JB = panel width (32 or 64).
template <typename T, int JB>
__global__ void __launch_bounds__(256) trsm_kernel(T *__restrict__ A, int lda, int c0, int c1, int k)
{
    main_stream_priority();
    constexpr int KQ = JB / 4, KP = KQ + 2, RS = 4 * KP + 1, NL = JB * JB / 256;
    __shared__ T Lq[JB][RS];
    const int lane = threadIdx.x & 63, p = lane & 3;
    const int wcol0 = c0 + (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    const int col = wcol0 + (lane >> 2);
    const bool live = col < c1;
    // the column's rows and L11 are loaded together (all loads in flight before the first is
    // used; filling LDS load by load had cost two dependent memory latencies more)
    T x[KQ], lv[NL];
#pragma unroll
    for (int q = 0; q < KQ; ++q)
        x[q] = live ? A[(size_t)(k + 4 * q + p) * lda + col] : T(0);
#pragma unroll
    for (int r = 0; r < NL; ++r)
    {
        const int idx = threadIdx.x + 256 * r, i = idx / JB, t = idx % JB;
        lv[r] = t < i ? A[(size_t)(k + i) * lda + k + t] : T(0);
    }
#pragma unroll
    for (int r = 0; r < NL; ++r)
    {
        const int idx = threadIdx.x + 256 * r, i = idx / JB, t = idx % JB;
        Lq[t][(i & 3) * KP + (i >> 2)] = lv[r];
    }
    __syncthreads();
    if (wcol0 >= c1)
        return;
#pragma unroll
    for (int t = 0; t < JB - 1; ++t)
    {
        const int kt = t >> 2, pt = t & 3;
        T xt;
        switch (pt)
        {
        case 0: xt = quad_bcast<0>(x[kt]); break;
        case 1: xt = quad_bcast<1>(x[kt]); break;
        case 2: xt = quad_bcast<2>(x[kt]); break;
        default: xt = quad_bcast<3>(x[kt]); break;
        }
        const T *lrow = &Lq[t][p * KP];
        {
            const T v = fma(-lrow[kt], xt, x[kt]);
            if (p > pt)
                x[kt] = v;
        }
#pragma unroll
        for (int q = kt + 1; q < KQ; ++q)
            x[q] = fma(-lrow[q], xt, x[q]);
    }
    if (live)
    {
#pragma unroll
        for (int q = 0; q < KQ; ++q)
            A[(size_t)(k + 4 * q + p) * lda + col] = x[q];
    }
}

// A22 -= L21 U12 on MFMA.  64x64 output tile per workgroup, 32x32 per wave (2x2 MFMA
// 16x16 tiles).  Operand lane layout of the 16x16x4 MFMAs: lane l supplies A[l%16][l/16]
// and B[l/16][l%16]; the accumulator row of element v is Mfma<T>::out_row.  Each 64-wide K
// segment (one panel's L21 U12) is summed from zero and subtracted from C once (as dgemm's
// C - A B: one rounding at |C|'s magnitude per panel, not one per k on diagonally dominant
// updates); the C tile stays in registers from the start to the end, across segments.  K is
// staged through LDS in chunks (16 wide for FP64, 32 for FP32), software-pipelined: the next
// chunk's operands are loaded into registers while the current chunk feeds the MFMAs, so
// global latency overlaps the matrix work (at most 128 VGPRs: four workgroups per CU).
// Tile order (1-D grid): consecutive workgroups go round-robin to the 8 XCDs, each with its
// own L2, so workgroup g works for XCD g % 8, which owns a stripe of cpx column tiles and
// sweeps it row tile by row tile: its U12 stripe stays in its L2 and each L21 tile is
// reused across the stripe while hot (rather than every XCD streaming all of U12 per row).
constexpr int kXcds = 8;
// C = rows [row0, row0 + m2) x columns [col0, col0 + n2) of A; the update subtracts
// A[C rows, kk0 .. kk0 + K) times A[kk0 .. kk0 + K, C columns].
// 16-byte vector of T (one global_load_dwordx4 / ds_write_b128)
template <typename T> struct Vec16;
template <> struct Vec16<double>
{
    typedef double type __attribute__((ext_vector_type(2)));
};
template <> struct Vec16<float>
{
    typedef float type __attribute__((ext_vector_type(4)));
};
template <typename T> using V16 = typename Vec16<T>::type;

template <typename T, bool AL>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) gemm_kernel(T *__restrict__ A, int lda, int row0, int m2, int col0, int n2,
                                                   int kk0, int K, int cpx)
{
    main_stream_priority<2>();
    constexpr int BM = 64, BN = 64, KC = sizeof(T) == 8 ? 16 : 32;
    // Row strides 36 and 80 elements: the operand reads (16 rows x 4 k, 4 k x 16 columns per
    // wave) then fall on distinct LDS banks apart from the two passes a 64-lane read needs.
    __shared__ __align__(16) T As[BM][KC + 4];
    __shared__ __align__(16) T Bs[KC][BN + 16];
    const int tid = threadIdx.x;
    const int xcd = blockIdx.x % kXcds, li = blockIdx.x / kXcds;
    const int tr = li / cpx, tc = xcd * cpx + li % cpx;
    if (tr * BM >= m2 || tc * BN >= n2)
        return; // the whole workgroup: outside C
    const int i0 = tr * BM, j0 = tc * BN;
    const int rowA = row0, colB = col0, k = kk0, jb = K;
    using acc_t = typename Mfma<T>::acc_t;
    const int w = tid >> 6, lane = tid & 63;
    const int wm = w >> 1, wn = w & 1;
    const int lr = lane & 15, lk = lane >> 4;
    T cprev[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int r = i0 + 32 * wm + 16 * a + Mfma<T>::out_row(lk, v);
                const int c = j0 + 32 * wn + 16 * b + lr;
                cprev[a][b][v] = (r < m2 && c < n2) ? A[(size_t)(rowA + r) * lda + colB + c] : T(0);
            }
    acc_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            acc[a][b] = acc_t{0, 0, 0, 0};
    // software pipeline: chunk k0 + KC is fetched into registers while chunk k0 is multiplied.
    // FP32 with an aligned lda (VL) loads 16 bytes per access (element loads at the matrix
    // edges); FP64 loads elements (its 16-byte variant measured slower).
    constexpr bool VL = AL && sizeof(T) == 4;
    constexpr int VEC = VL ? 16 / (int)sizeof(T) : 1;
    using R = std::conditional_t<VL, V16<T>, T>;
    constexpr int PA = BM * KC / 256 / VEC, PB = KC * BN / 256 / VEC;
    R ra[PA], rb[PB];
    auto fetch = [&](int c0) {
#pragma unroll
        for (int i = 0; i < PA; ++i)
        {
            const int idx = tid + 256 * i, r = idx / (KC / VEC), t = (idx % (KC / VEC)) * VEC;
            const T *src = A + (size_t)(rowA + i0 + r) * lda + k + c0 + t;
            if constexpr (VL)
            {
                if (i0 + r < m2 && c0 + t + VEC <= jb)
                    ra[i] = *reinterpret_cast<const R *>(src);
                else
#pragma unroll
                    for (int e = 0; e < VEC; ++e)
                        ra[i][e] = (i0 + r < m2 && c0 + t + e < jb) ? src[e] : T(0);
            }
            else
                ra[i] = (i0 + r < m2 && c0 + t < jb) ? *src : T(0);
        }
#pragma unroll
        for (int i = 0; i < PB; ++i)
        {
            const int idx = tid + 256 * i, t = idx / (BN / VEC), c = (idx % (BN / VEC)) * VEC;
            const T *src = A + (size_t)(k + c0 + t) * lda + colB + j0 + c;
            if constexpr (VL)
            {
                if (c0 + t < jb && j0 + c + VEC <= n2)
                    rb[i] = *reinterpret_cast<const R *>(src);
                else
#pragma unroll
                    for (int e = 0; e < VEC; ++e)
                        rb[i][e] = (c0 + t < jb && j0 + c + e < n2) ? src[e] : T(0);
            }
            else
                rb[i] = (c0 + t < jb && j0 + c < n2) ? *src : T(0);
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < jb; k0 += KC)
    {
        if (k0)
            __syncthreads(); // the previous chunk's operands are consumed
#pragma unroll
        for (int i = 0; i < PA; ++i)
        {
            const int idx = tid + 256 * i;
            *reinterpret_cast<R *>(&As[idx / (KC / VEC)][(idx % (KC / VEC)) * VEC]) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < PB; ++i)
        {
            const int idx = tid + 256 * i;
            *reinterpret_cast<R *>(&Bs[idx / (BN / VEC)][(idx % (BN / VEC)) * VEC]) = rb[i];
        }
        __syncthreads();
        if (k0 + KC < jb)
            fetch(k0 + KC);
        const int kend = min(KC, (jb - k0 + 3) & ~3);
        for (int kk = 0; kk < kend; kk += 4)
        {
            T af[2], bf[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
                af[a] = As[32 * wm + 16 * a + lr][kk + lk];
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bf[b] = Bs[kk + lk][32 * wn + 16 * b + lr];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = Mfma<T>::mma(af[a], bf[b], acc[a][b]);
        }
        if ((k0 + KC) % NB == 0 || k0 + KC >= jb)
        {
            // end of a 64-wide K segment (one panel): subtract its sum, as a per-panel update
            // would, so the result does not depend on how many panels one launch covers
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                {
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        cprev[a][b][v] = cprev[a][b][v] - acc[a][b][v];
                    acc[a][b] = acc_t{0, 0, 0, 0};
                }
        }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int r = i0 + 32 * wm + 16 * a + Mfma<T>::out_row(lk, v);
                const int c = j0 + 32 * wn + 16 * b + lr;
                if (r < m2 && c < n2)
                    A[(size_t)(rowA + r) * lda + colB + c] = cprev[a][b][v];
            }
}

// Gather rows: Y[i] = B[perm[i]] (row-major, nrhs columns).
template <typename T>
__global__ void __launch_bounds__(256) permute_rows_kernel(const T *__restrict__ B, T *__restrict__ Y, int n, int nrhs,
                                                           const int *__restrict__ perm)
{
    main_stream_priority();
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= n * nrhs)
        return;
    const int i = idx / nrhs, q = idx - i * nrhs;
    Y[idx] = B[(size_t)perm[i] * nrhs + q];
}

// Inverses of the BS x BS diagonal blocks of the FP32 factors (identity padding past n), for
// the single-launch solves (inside FP64 refinement): out[(2 ib + 0)] = L_bb^-1 (unit lower),
// out[(2 ib + 1)] = U_bb^-1, row-major BS x BS each.  Thread j forms column j by substitution.
template <int BS>
__global__ void __launch_bounds__(256) diag_inverse_kernel(const float *__restrict__ A, int lda, int n,
                                                           float *__restrict__ out)
{
    main_stream_priority();
    __shared__ float M[BS][BS + 1];
    __shared__ float X[BS][BS + 1];
    const int ib = blockIdx.x, upper = blockIdx.y, tid = threadIdx.x;
    const int r0 = ib * BS, nr = min(BS, n - r0);
    for (int idx = tid; idx < BS * BS; idx += 256)
    {
        const int r = idx / BS, c = idx % BS;
        M[r][c] = (r < nr && c < nr) ? A[(size_t)(r0 + r) * lda + r0 + c] : (r == c ? 1.0f : 0.0f);
        X[r][c] = 0.0f;
    }
    __syncthreads();
    if (tid < BS)
    {
        const int j = tid;
        if (!upper)
        {
            X[j][j] = 1.0f;
            for (int i = j + 1; i < BS; ++i)
            {
                float acc = 0.0f;
                for (int k = j; k < i; ++k)
                    acc = fmaf(M[i][k], X[k][j], acc);
                X[i][j] = -acc;
            }
        }
        else
        {
            X[j][j] = 1.0f / M[j][j];
            for (int i = j - 1; i >= 0; --i)
            {
                float acc = 0.0f;
                for (int k = i + 1; k <= j; ++k)
                    acc = fmaf(M[i][k], X[k][j], acc);
                X[i][j] = -acc / M[i][i];
            }
        }
    }
    __syncthreads();
    float *o = out + (size_t)(2 * ib + upper) * BS * BS;
    for (int idx = tid; idx < BS * BS; idx += 256)
        o[idx] = X[idx / BS][idx % BS];
}

// Block triangular solve on B (n x nrhs, row-major), in place, in ONE launch of one
// workgroup per BS-row block (64 rows for FP64, 128 for FP32).  LOWER: unit lower L, forward order; else upper U, backward
// order.  Workgroups take solve-order indices from a ticket counter, so every smaller index
// belongs to a workgroup already running; workgroup t subtracts the contributions of blocks
// 0 .. t-1 in that order as their flags (sync[kFlowHeader + k] == epoch) appear, each block
// as a dot product subtracted once (as a blocked dtrsv with a GEMV update does; folding each
// term into the running value loses accuracy over hundreds of blocks), then solves its own
// block and raises its flag.  The diagonal solve is, for FP32, the product with the block's
// precomputed inverse; for FP64, substitution in LAPACK dtrsv's order (column sweep, lane =
// row).  Waits are bounded: a timed-out wait sets *err and the workgroup writes NaN.
constexpr int kFlowHeader = 64; // sync[0], sync[1]: tickets of the lower and upper solves

__device__ __forceinline__ bool flow_wait(const unsigned *flag, unsigned epoch, int *err)
{
    // relaxed polling (device-coherent loads, no cache invalidation per poll), one acquire
    // fence once the flag is seen
    unsigned spins = 0;
    bool ok = true;
    while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != epoch)
    {
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit ||
            ((spins & 1023u) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
        {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = false;
            break;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return ok;
}

__global__ void flow_reset_kernel(unsigned *__restrict__ sync, int *__restrict__ err)
{
    main_stream_priority();
    if (threadIdx.x < 2)
        sync[threadIdx.x] = 0u;
    if (threadIdx.x == 2)
        *err = 0;
}

constexpr int kXs = kMaxRhs + 4; // row stride of the right-hand-side tiles (16-byte rows)
constexpr int kFlowBs32 = 128;   // block rows of the FP32 solves (and of their diagonal inverses)

// out(r, q, s) for s = sum_c M[r][c] X[c][q], the BS x (columns < 4 nq4) block product on
// MFMA: wave w takes rows 16w .. 16w + 15 and every 16-column tile, K in steps of 4 from
// c = 0 (the order within a step is the matrix core's).  X columns at or beyond 4 nq4 enter
// as zeros and their outputs are not reported.
template <typename T, int BS, int NT, typename F>
__device__ __forceinline__ void block_product_mfma(const T (*M)[BS + 1], const T (*X)[kXs], int nq4, F &&out)
{
    static_assert(BS == 16 * (NT / 64), "one 16-row tile per wave");
    using acc_t = typename Mfma<T>::acc_t;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int ncol = 4 * nq4, nct = (ncol + 15) / 16;
    for (int ct = 0; ct < nct; ++ct)
    {
        const int q = 16 * ct + lr;
        acc_t c = acc_t{0, 0, 0, 0};
#pragma unroll 8
        for (int kk = 0; kk < BS; kk += 4)
        {
            const T a = M[16 * w + lr][kk + lk];
            const T b = q < ncol ? X[kk + lk][q] : T(0);
            c = Mfma<T>::mma(a, b, c);
        }
        if (q < ncol)
        {
#pragma unroll
            for (int v = 0; v < 4; ++v)
                out(16 * w + Mfma<T>::out_row(lk, v), q, c[v]);
        }
    }
}

// Substitution with the diagonal block of one block row (LAPACK dtrsv order): lane = row,
// NQ right-hand sides per wave in registers; row k's value is broadcast with readlane.  The
// lane's row of the block is read from LDS eight columns at a time, the next eight while the
// current ones are used, so a step waits on no LDS access: the chain per step is the FMA,
// the broadcast and (upper) the division on the pivot lane.
template <typename T, bool LOWER, int BS, int NQ, int QW>
__device__ __forceinline__ void flow_substitute(T (&x)[QW], const T (*Ls)[BS + 1], int lane, T diag)
{
    constexpr int KC = 8;
    T lc[KC], ln[KC];
    // upper: the lane's own reciprocal 1/u_kk, correctly rounded (a true division), ahead of
    // the chain; a step's quotient is then q0 = x r, e = x - u q0 (exact with the FMA),
    // q = q0 + e r: the correctly rounded x / u_kk (Markstein's correction with a correctly
    // rounded reciprocal; exact wherever e does not underflow, |x| >= 2^-969, else within an
    // ulp) in three dependent operations instead of a division sequence
    const T rdiag = LOWER ? T(0) : T(1) / diag;
#pragma unroll
    for (int u = 0; u < KC; ++u)
        lc[u] = Ls[lane][LOWER ? u : BS - 1 - u];
    for (int kc = 0; kc < BS; kc += KC)
    {
        if (kc + KC < BS)
        {
#pragma unroll
            for (int u = 0; u < KC; ++u)
                ln[u] = Ls[lane][LOWER ? kc + KC + u : BS - 1 - (kc + KC + u)];
        }
#pragma unroll
        for (int u = 0; u < KC; ++u)
        {
            const int k = LOWER ? kc + u : BS - 1 - (kc + u);
#pragma unroll
            for (int j = 0; j < NQ; ++j)
            {
                if (!LOWER)
                {
                    // every lane forms the quotient and the pivot lane keeps it: no branch, so
                    // the right-hand sides' quotients interleave
                    const T q0 = x[j] * rdiag;
                    const T e = fma(-q0, diag, x[j]);
                    const T q = fma(e, rdiag, q0);
                    x[j] = lane == k ? q : x[j];
                }
                const T xk = bcast_lane(x[j], k);
                if (LOWER ? lane > k : lane < k)
                    x[j] = fma(-lc[u], xk, x[j]);
            }
        }
#pragma unroll
        for (int u = 0; u < KC; ++u)
            lc[u] = ln[u];
    }
}

template <typename T, bool LOWER, int BS, int NT>
__global__ void __launch_bounds__(NT) trsv_flow_kernel(const T *__restrict__ A, int lda, int n, T *__restrict__ B,
                                                        int nrhs, const float *__restrict__ dinv,
                                                        unsigned *__restrict__ sync, unsigned epoch,
                                                        int *__restrict__ err)
{
    main_stream_priority<3>();
    __shared__ T Ls[BS][BS + 1];
    __shared__ __align__(16) T xs[BS][kXs];
    __shared__ __align__(16) T acc[BS][kXs];
    __shared__ int s_t, s_ok;
    const int tid = threadIdx.x;
    if (tid == 0)
    {
        s_t = (int)__hip_atomic_fetch_add(&sync[LOWER ? 0 : 1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_ok = 1;
    }
    __syncthreads();
    unsigned *flag = sync + kFlowHeader;
    const int nblk = (n + BS - 1) / BS;
    const int t = s_t; // solve order
    const int ib = LOWER ? t : nblk - 1 - t;
    const int r0 = ib * BS, nr = min(BS, n - r0);
    const int nq4 = (nrhs + 3) / 4, nqp = 4 * nq4;
    constexpr int kPer = BS * BS / NT;
    // the diagonal block's operator (inverse for FP32, the factor block for FP64) and the
    // first off-diagonal block go to registers now; later blocks are fetched one ahead and
    // staged in LDS before the wait for their right-hand side
    T dreg[kPer], lreg[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k)
    {
        const int idx = tid + NT * k, r = idx / BS, c = idx % BS;
        if constexpr (sizeof(T) == 4)
            dreg[k] = dinv[(size_t)(2 * ib + (LOWER ? 0 : 1)) * BS * BS + idx];
        else
            dreg[k] = (r < nr && c < nr) ? A[(size_t)(r0 + r) * lda + r0 + c] : (r == c ? T(1) : T(0));
    }
    auto fetch_l = [&](int k) {
        const int c0 = (LOWER ? k : nblk - 1 - k) * BS, nc = min(BS, n - c0);
#pragma unroll
        for (int kk = 0; kk < kPer; ++kk)
        {
            const int idx = tid + NT * kk, r = idx / BS, c = idx % BS;
            lreg[kk] = (r < nr && c < nc) ? A[(size_t)(r0 + r) * lda + c0 + c] : T(0);
        }
    };
    auto stage = [&](const T *reg) {
#pragma unroll
        for (int kk = 0; kk < kPer; ++kk)
        {
            const int idx = tid + NT * kk;
            Ls[idx / BS][idx % BS] = reg[kk];
        }
    };
    for (int idx = tid; idx < BS * nqp; idx += NT)
    {
        const int r = idx / nqp, q = idx - r * nqp;
        acc[r][q] = (r < nr && q < nrhs) ? B[(size_t)(r0 + r) * nrhs + q] : T(0);
    }
    if (t > 0)
    {
        fetch_l(0);
        stage(lreg);
        if (t > 1)
            fetch_l(1);
    }
    for (int k = 0; k < t; ++k)
    {
        const int c0 = (LOWER ? k : nblk - 1 - k) * BS, nc = min(BS, n - c0);
        if (tid == 0 && !flow_wait(&flag[k], epoch, err))
            s_ok = 0;
        __syncthreads();
        for (int idx = tid; idx < BS * nqp; idx += NT)
        {
            const int r = idx / nqp, q = idx - r * nqp;
            xs[r][q] = (r < nc && q < nrhs) ? xload(&B[(size_t)(c0 + r) * nrhs + q]) : T(0);
        }
        __syncthreads();
        block_product_mfma<T, BS, NT>(Ls, xs, nq4, [&](int r, int q, T a) { acc[r][q] -= a; });
        __syncthreads();
        if (k + 1 < t)
        {
            stage(lreg);
            if (k + 2 < t)
                fetch_l(k + 2);
        }
    }
    stage(dreg);
    __syncthreads();
    const T poison = s_ok ? T(0) : T(NAN);
    if constexpr (sizeof(T) == 4)
    {
        block_product_mfma<T, BS, NT>(Ls, acc, nq4, [&](int r, int q, T a) {
            if (r < nr && q < nrhs)
                xstore(&B[(size_t)(r0 + r) * nrhs + q], a + poison);
        });
    }
    else
    {
        // wave wv holds the right-hand sides wv, wv + 4, ... (at most kMaxRhs / 4) in
        // registers and advances them together (flow_substitute).
        static_assert(BS == 64 && NT == 256, "FP64 substitution: one row per lane, four waves");
        const int lane = tid & 63, wv = tid >> 6;
        constexpr int QW = kMaxRhs / 4;
        T x[QW];
#pragma unroll
        for (int j = 0; j < QW; ++j)
        {
            const int q = wv + 4 * j;
            x[j] = q < nrhs ? acc[lane][q] : T(0);
        }
        const int nq = (nrhs - wv + 3) / 4; // right-hand sides of this wave (uniform)
        const T lrow_diag = Ls[lane][lane];
        // the count of right-hand sides is a template argument: a runtime bound inside the
        // unrolled steps turns into a branch cascade that copies every x at every exit
        switch (nq)
        {
        case 1: flow_substitute<T, LOWER, BS, 1>(x, Ls, lane, lrow_diag); break;
        case 2: flow_substitute<T, LOWER, BS, 2>(x, Ls, lane, lrow_diag); break;
        case 3: flow_substitute<T, LOWER, BS, 3>(x, Ls, lane, lrow_diag); break;
        case 4: flow_substitute<T, LOWER, BS, 4>(x, Ls, lane, lrow_diag); break;
        case 5: flow_substitute<T, LOWER, BS, 5>(x, Ls, lane, lrow_diag); break;
        case 6: flow_substitute<T, LOWER, BS, 6>(x, Ls, lane, lrow_diag); break;
        case 7: flow_substitute<T, LOWER, BS, 7>(x, Ls, lane, lrow_diag); break;
        case 8: flow_substitute<T, LOWER, BS, 8>(x, Ls, lane, lrow_diag); break;
        default: break;
        }
#pragma unroll
        for (int j = 0; j < QW; ++j)
        {
            const int q = wv + 4 * j;
            if (q < nrhs && lane < nr)
                xstore(&B[(size_t)(r0 + lane) * nrhs + q], x[j] + poison);
        }
    }
    // every wave's solution stores (device-coherent) complete; thread 0's release store of
    // the flag covers the workgroup
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (tid == 0)
        __hip_atomic_store(&flag[t], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// s4[j] += sum over c in [c0, c1), c = lane mod 64, of arow[j][c] * x[c], for four factor rows:
// each lane walks its columns in increasing order (the association of a lane-strided dot
// product); the loads of four columns per row are issued before their FMAs, so that 16 of a
// lane's loads are in flight at once (the single-workgroup solves are bound by how many bytes
// one CU keeps in flight, not by arithmetic).
template <typename T>
__device__ __forceinline__ void dot4_lane_strided(const T *const (&arow)[4], const T *__restrict__ x, int c0, int c1,
                                                  int lane, T (&s4)[4])
{
    constexpr int U = 4;
    for (int cb = c0 + lane; cb < c1; cb += 64 * U)
    {
        T av[U][4], xv[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            const int c = cb + 64 * u;
            const bool ok = c < c1;
            const int cc = ok ? c : cb; // cb < c1: an in-range column
#pragma unroll
            for (int j = 0; j < 4; ++j)
                av[u][j] = arow[j][cc];
            xv[u] = ok ? x[cc] : T(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            if (cb + 64 * u < c1)
            {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    s4[j] = fma(av[u][j], xv[u], s4[j]);
            }
        }
    }
}

// Whole triangular solve in ONE workgroup (512 threads, at most 96 VGPRs: it fits on a CU
// beside the side stream's point-velocity waves) for small systems: x (n x nrhs) lives in LDS; blocks of 64
// rows are taken in solve order, each first reduced by the already solved part (dot products
// over contiguous factor rows, 8 rows per wave) and then
// solved with its diagonal block by substitution (wave 0, LAPACK dtrsv order).  Used when
// the LDS holds x (n * nrhs * sizeof(T) <= kTrsvWgBytes).
constexpr int kTrsvWgBytes = 64 * 1024;
constexpr int kTrsvWgThreads = 512;
template <typename T, bool LOWER>
__global__ void __launch_bounds__(kTrsvWgThreads) __attribute__((amdgpu_waves_per_eu(5))) trsv_wg_kernel(const T *__restrict__ A, int lda, int n, T *__restrict__ B,
                                                      int nrhs)
{
    main_stream_priority<3>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *x = (T *)smem_raw; // n x nrhs
    __shared__ T Ds[NB][NB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int i = tid; i < n * nrhs; i += kTrsvWgThreads)
        x[i] = B[i];
    __syncthreads();
    const int nblk = (n + NB - 1) / NB;
    for (int t = 0; t < nblk; ++t)
    {
        const int ib = LOWER ? t : nblk - 1 - t;
        const int r0 = ib * NB, nr = min(NB, n - r0);
        // solved columns: [0, r0) for L, [r0 + nr, n) for U
        const int c0 = LOWER ? 0 : r0 + nr, c1 = LOWER ? r0 : n;
        // 8 waves x 2 passes x 4 rows = the 64 rows of the block, each pass's four rows' loads
        // in flight together
        for (int pass = 0; pass < 2; ++pass)
        {
            const int ra = 8 * wv + 4 * pass;
            const T *arow[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                arow[j] = A + (size_t)(r0 + min(ra + j, nr - 1)) * lda;
            for (int q = 0; q < nrhs; ++q)
            {
                T s4[4] = {T(0), T(0), T(0), T(0)};
#pragma unroll 4
                for (int c = c0 + lane; c < c1; c += 64)
                {
                    const T xv = x[c * nrhs + q];
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        s4[j] = fma(arow[j][c], xv, s4[j]);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j)
                {
                    T s2 = s4[j];
                    for (int off = 32; off > 0; off >>= 1)
                        s2 += __shfl_xor(s2, off);
                    if (lane == 0 && ra + j < nr)
                        x[(r0 + ra + j) * nrhs + q] -= s2;
                }
            }
        }
        __syncthreads();
        // the diagonal block, staged in LDS by all threads (identity padding past n)
        for (int idx = tid; idx < NB * NB; idx += kTrsvWgThreads)
        {
            const int r = idx / NB, c = idx % NB;
            Ds[r][c] = (r < nr && c < nr) ? A[(size_t)(r0 + r) * lda + r0 + c] : (r == c ? T(1) : T(0));
        }
        __syncthreads();
        if (wv == 0)
        {
            const T dd = Ds[lane][lane];
            for (int q = 0; q < nrhs; ++q)
            {
                T v = lane < nr ? x[(r0 + lane) * nrhs + q] : T(0);
                for (int kk = 0; kk < NB; ++kk)
                {
                    const int k = LOWER ? kk : NB - 1 - kk;
                    if (!LOWER && lane == k)
                        v = v / dd;
                    const T vk = bcast_lane(v, k);
                    if (LOWER ? lane > k : lane < k)
                        v = fma(-Ds[lane][k], vk, v);
                }
                if (lane < nr)
                    x[(r0 + lane) * nrhs + q] = v;
            }
        }
        __syncthreads();
    }
    for (int i = tid; i < n * nrhs; i += kTrsvWgThreads)
        B[i] = x[i];
}

// Single right-hand side, whole triangular solve in ONE workgroup with look-ahead (the time
// loop's solve).  Blocks of BS rows are solved in order; block s's reduction by the solved
// part is split into the blocks solved up to s - 2 ("old", a dot product over contiguous
// factor rows, as in trsv_wg_kernel) and block s - 1 ("new").  While wave 0 solves block s
// (the new part from the staged off-diagonal block, then substitution with the staged
// diagonal block, LAPACK dtrsv order), waves 1-7 stage block s + 1's diagonal block and its
// off-diagonal block against block s, and form its old part: only wave 0's short chain is
// on the critical path.  Each row is reduced once by old + new (one subtraction per block,
// as a blocked dtrsv).  BS = 32 keeps the LDS (x and four BS x (BS + 1) tiles: 42 KB for
// n = 1,024 in FP64) small enough to share a CU with the side stream's workgroups.
constexpr int kTrsvLaThreads = 512;
constexpr int kTrsvLaBs = 32;
constexpr size_t kTrsvLaMaxX = 32 * 1024; // x in LDS: n <= 4,096 (FP64)
template <typename T, bool LOWER, int BS>
__global__ void __launch_bounds__(kTrsvLaThreads) __attribute__((amdgpu_waves_per_eu(5)))
    trsv_la_kernel(const T *__restrict__ A, int lda, int n, T *__restrict__ b)
{
    main_stream_priority<3>();
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *x = (T *)smem_raw; // n
    __shared__ T Dt[2][BS][BS + 1], Ot[2][BS][BS + 1], sold[2][BS];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nb = (n + BS - 1) / BS;
    auto blk = [&](int s) { return LOWER ? s : nb - 1 - s; };
    for (int i = tid; i < n; i += kTrsvLaThreads)
        x[i] = b[i];
    // helpers' work for solve step s (block blk(s)), done during step s - 1 (or up front for
    // s = 0): stage the diagonal block and the off-diagonal block against blk(s - 1), and the
    // old part over the blocks solved at steps <= s - 2
    auto prepare = [&](int s, int htid, int hw, int nhw) {
        const int ib = blk(s), r0 = ib * BS, nr = min(BS, n - r0);
        const int buf = s & 1;
        for (int idx = htid; idx < BS * BS; idx += nhw * 64)
        {
            const int r = idx / BS, c = idx % BS;
            Dt[buf][r][c] = (r < nr && c < nr) ? A[(size_t)(r0 + r) * lda + r0 + c] : (r == c ? T(1) : T(0));
        }
        if (s >= 1)
        {
            const int p0 = blk(s - 1) * BS, np = min(BS, n - p0);
            for (int idx = htid; idx < BS * BS; idx += nhw * 64)
            {
                const int r = idx / BS, c = idx % BS;
                Ot[buf][r][c] = (r < nr && c < np) ? A[(size_t)(r0 + r) * lda + p0 + c] : T(0);
            }
        }
        // old columns: blocks solved at steps <= s - 2
        int c0 = 0, c1 = 0;
        if (s >= 2)
        {
            if (LOWER)
                c0 = 0, c1 = (s - 1) * BS;
            else
                c0 = (nb - s + 1) * BS, c1 = n;
        }
        for (int g = hw; g < BS / 4; g += nhw)
        {
            const int ra = 4 * g;
            const T *arow[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                arow[j] = A + (size_t)(r0 + min(ra + j, nr - 1)) * lda;
            T s4[4] = {T(0), T(0), T(0), T(0)};
            dot4_lane_strided<T>(arow, x, c0, c1, lane, s4);
#pragma unroll
            for (int j = 0; j < 4; ++j)
            {
                T s2 = s4[j];
                for (int off = 32; off > 0; off >>= 1)
                    s2 += __shfl_xor(s2, off);
                if (lane == 0)
                    sold[buf][ra + j] = s2;
            }
        }
    };
    __syncthreads();
    prepare(0, tid, wv, 8);
    __syncthreads();
    for (int s = 0; s < nb; ++s)
    {
        const int buf = s & 1;
        if (wv == 0)
        {
            const int ib = blk(s), r0 = ib * BS, nr = min(BS, n - r0);
            const int r = lane < BS ? lane : BS - 1;
            T snew = T(0);
            if (s >= 1)
            {
                const int p0 = blk(s - 1) * BS;
                T a0 = T(0), a1 = T(0);
#pragma unroll 4
                for (int k = 0; k < BS; k += 2)
                {
                    a0 = fma(Ot[buf][r][k], x[min(p0 + k, n - 1)], a0);
                    a1 = fma(Ot[buf][r][k + 1], x[min(p0 + k + 1, n - 1)], a1);
                }
                snew = a0 + a1;
            }
            T v = lane < nr ? x[r0 + lane] - (sold[buf][r] + snew) : T(0);
            const T dd = Dt[buf][r][r];
            for (int kk = 0; kk < BS; ++kk)
            {
                const int k = LOWER ? kk : BS - 1 - kk;
                if (!LOWER && lane == k)
                    v = v / dd;
                const T vk = bcast_lane(v, k);
                if (LOWER ? (lane > k && lane < BS) : lane < k)
                    v = fma(-Dt[buf][r][k], vk, v);
            }
            if (lane < nr)
                x[r0 + lane] = v;
        }
        else if (s + 1 < nb)
        {
            prepare(s + 1, tid - 64, wv - 1, 7);
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += kTrsvLaThreads)
        b[i] = x[i];
}

// R = B - A X on FP64 MFMA (v_mfma_f64_16x16x4f64), for the refinement's FP64 residuals: a
// wave forms a 16-row x 16-column tile of A X (QT tiles side by side for nrhs <= 16 QT) over
// one K slice; residual_sum_kernel adds the slices in order.  Within a slice the K order is
// permuted so that each lane streams 4 consecutive elements of its row of A: at MFMA step j
// of a 16-wide K step, lane (r, c) supplies A[r][k + 4c + j] and X[k + 4c + j][q] (any
// fixed order of the sum is a valid residual).
template <int QT>
__global__ void __launch_bounds__(256) residual_mfma_kernel(const double *__restrict__ A, int lda, int n,
                                                            const double *__restrict__ X, int nrhs, int kslice,
                                                            double *__restrict__ part)
{
    main_stream_priority();
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    const int r0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (r0 >= n)
        return;
    const int s = blockIdx.y, kb = s * kslice, ke = min(n, kb + kslice);
    const int row = r0 + lr;
    const double *a = A + (size_t)(row < n ? row : n - 1) * lda;
    double4_t acc[QT];
#pragma unroll
    for (int t = 0; t < QT; ++t)
        acc[t] = double4_t{0, 0, 0, 0};
    for (int k = kb; k < ke; k += 16)
    {
        double av[4];
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int kk = k + 4 * lk + j;
            av[j] = (row < n && kk < ke) ? a[kk] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
        {
            const int kk = k + 4 * lk + j;
#pragma unroll
            for (int t = 0; t < QT; ++t)
            {
                const int q = 16 * t + lr;
                const double bv = (kk < ke && q < nrhs) ? X[(size_t)kk * nrhs + q] : 0.0;
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[j], bv, acc[t], 0, 0, 0);
            }
        }
    }
    // element v of lane (c, r) is (A X)[r0 + 4v + c][16t + r]
#pragma unroll
    for (int t = 0; t < QT; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int rr = r0 + 4 * v + lk, q = 16 * t + lr;
            if (rr < n && q < nrhs)
                part[((size_t)s * n + rr) * nrhs + q] = acc[t][v];
        }
}

__global__ void __launch_bounds__(256) residual_sum_kernel(const double *__restrict__ part, int slices, size_t count,
                                                           const double *__restrict__ B, double *__restrict__ R)
{
    main_stream_priority();
    const size_t i = blockIdx.x * (size_t)256 + threadIdx.x;
    if (i >= count)
        return;
    double ax = part[i];
    for (int s = 1; s < slices; ++s)
        ax = ax + part[(size_t)s * count + i];
    R[i] = B[i] - ax;
}

__global__ void f64_to_f32_kernel(const double *__restrict__ a, float *__restrict__ b, size_t count)
{
    main_stream_priority();
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256)
        b[i] = (float)a[i];
}

// Per-column max-norms of D and X (n x nrhs) -> out[0:nrhs] = |D|inf, out[nrhs:2nrhs] = |X|inf.
// out[q] = max_i |M[i][q]| (M: n x nrhs row-major); NaN counts as infinite.
__global__ void __launch_bounds__(256) colabsmax_kernel(const double *__restrict__ M, int n, int nrhs,
                                                        double *__restrict__ out)
{
    main_stream_priority();
    const int q = blockIdx.x;
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256)
    {
        const double v = fabs(M[(size_t)i * nrhs + q]);
        m = v <= DBL_MAX ? fmax(m, v) : INFINITY;
    }
    __shared__ double sm[256];
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1)
    {
        if (threadIdx.x < s)
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[q] = sm[0];
}

// X += (double) D32
__global__ void add_f32_kernel(double *__restrict__ X, const float *__restrict__ D, size_t count)
{
    main_stream_priority();
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < count; i += (size_t)gridDim.x * 256)
        X[i] = X[i] + (double)D[i];
}

// ---------------------------------------------------------------------------------------
// Host side
struct Workspace
{
    void *panelx = nullptr;
    int *err = nullptr; // [0]: first zero pivot, [1]: a bounded wait timed out, [2]: no-pivot zero pivot,
                        // [3]: an interchange was due, [4]: workgroups done (unpivoted panel)
};
Workspace g_ws;
// rocBLAS handle for the outer trailing updates, created at the first factorisation that has
// one (n > 2048); bound to the library stream at every call.
rocblas_handle g_blas = nullptr;
// Pinned, coherent host copy of err[0..3], written by a kernel: the end-of-factorisation
// readback needs no blit (the runtime's copy kernels can wait for CU room behind the side
// stream's work).
volatile int *g_status_host = nullptr;
int *g_status_host_dev = nullptr;

__global__ void factor_reset_kernel(PanelRec *__restrict__ rec, int *__restrict__ status)
{
    main_stream_priority();
    // records tagged 1: never taken for rounds 0 and 1 (tag 0), and overwritten before the
    // rounds with tag 1 (2 and 3) read them
    for (int i = threadIdx.x; i < 2 * kMaxGroups; i += blockDim.x)
    {
        (&rec->slot[0][0])[i].key = 1ull << 63;
        (&rec->slot[0][0])[i].pos = 1u << 31;
    }
    if (threadIdx.x == 0)
    {
        status[0] = 0;        // first zero pivot
        status[1] = 0;        // a bounded wait timed out
        status[2] = 0x7F7F7F7F; // no-pivot zero pivots: min
        status[3] = 0;        // an interchange was due
    }
}

__global__ void status_to_host_kernel(const int *__restrict__ status, int *__restrict__ host)
{
    main_stream_priority();
    if (threadIdx.x < 4)
        host[threadIdx.x] = status[threadIdx.x];
}

// A timed-out wait of the single-launch solve (err[5]) is recorded in the pinned host word
// kAsyncErrorWord, where it stays until the host takes it at its next synchronisation.
constexpr int kAsyncErrorWord = 8;
__global__ void record_async_error_kernel(const int *__restrict__ err, int *__restrict__ host)
{
    main_stream_priority();
    if (threadIdx.x == 0 && *err)
        host[kAsyncErrorWord] = 1;
}

__global__ void identity_kernel(int *__restrict__ perm, int n)
{
    main_stream_priority();
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        perm[i] = i;
}

int ensure_ws()
{
    if (!g_ws.panelx)
    {
        PVL_HIP(hipMalloc(&g_ws.panelx, sizeof(PanelX<double>)));
        PVL_HIP(hipMalloc(&g_ws.err, 8 * sizeof(int)));
        PVL_HIP(hipMemset(g_ws.err, 0, 8 * sizeof(int)));
        void *h = nullptr;
        PVL_HIP(hipHostMalloc(&h, 16 * sizeof(int), hipHostMallocCoherent | hipHostMallocMapped));
        for (int i = 0; i < 16; ++i)
            ((int *)h)[i] = 0;
        g_status_host = (volatile int *)h;
        PVL_HIP(hipHostGetDevicePointer((void **)&g_status_host_dev, h, 0));
    }
    return PVL_OK;
}

} // namespace

int pvl::take_async_error()
{
    if (g_status_host && g_status_host[kAsyncErrorWord])
    {
        g_status_host[kAsyncErrorWord] = 0;
        set_error("a single-launch triangular solve timed out waiting for a block (its rows were written as NaN)");
        return PVL_ERR_HIP;
    }
    return PVL_OK;
}

void pvl::lu_release_workspace()
{
    if (g_blas)
        (void)rocblas_destroy_handle(g_blas);
    g_blas = nullptr;
    (void)hipFree(g_ws.panelx);
    (void)hipFree(g_ws.err);
    if (g_status_host)
        (void)hipHostFree((void *)g_status_host);
    g_ws = Workspace{};
    g_status_host = nullptr;
    g_status_host_dev = nullptr;
}

namespace
{

template <typename T> int panel_rows_max() { return 256; }

#define PVL_BLAS(x)                                                                                                   \
    do                                                                                                                \
    {                                                                                                                 \
        const rocblas_status s_ = (x);                                                                                \
        if (s_ != rocblas_status_success)                                                                             \
        {                                                                                                             \
            set_error("rocBLAS: %s (%s:%d)", rocblas_status_to_string(s_), __FILE__, __LINE__);                       \
            return PVL_ERR_HIP;                                                                                       \
        }                                                                                                             \
    } while (0)

// Configuration of a new handle: deterministic results (no split-K atomics), full-precision
// FP32 products, scalars from the host.
int blas_configure(rocblas_handle h)
{
    PVL_BLAS(rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed));
    PVL_BLAS(rocblas_set_math_mode(h, rocblas_default_math));
    PVL_BLAS(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
    return PVL_OK;
}

// The handle bound to the library stream; kept only once fully configured.
int blas_handle()
{
    if (!g_blas)
    {
        rocblas_handle h = nullptr;
        PVL_BLAS(rocblas_create_handle(&h));
        if (int st = blas_configure(h))
        {
            (void)rocblas_destroy_handle(h);
            return st;
        }
        g_blas = h;
    }
    PVL_BLAS(rocblas_set_stream(g_blas, stream()));
    return PVL_OK;
}

// Row-major C (m2 x n2, leading dimension lda) -= A (m2 x K) B (K x n2), all three inside the
// factorised matrix.  Column-major rocBLAS sees the transposes: C^T -= B^T A^T.
template <typename T> int blas_update(T *C, const T *Ap, const T *Bp, int lda, int m2, int n2, int K)
{
    if (int st = blas_handle())
        return st;
    const T alpha = T(-1), beta = T(1);
    if constexpr (sizeof(T) == 8)
        PVL_BLAS(rocblas_dgemm(g_blas, rocblas_operation_none, rocblas_operation_none, n2, m2, K, &alpha, Bp, lda,
                               Ap, lda, &beta, C, lda));
    else
        PVL_BLAS(rocblas_sgemm(g_blas, rocblas_operation_none, rocblas_operation_none, n2, m2, K, &alpha, Bp, lda,
                               Ap, lda, &beta, C, lda));
    return PVL_OK;
}

template <typename T> int factor_impl(int n, T *A, int *piv, int *info_out, bool pivoting, bool *would_pivot = nullptr)
{
    if (int st = ensure_ws())
        return st;
    static bool attr_set[2] = {false, false};
    const int rmax = panel_rows_max<T>();
    // one row beyond the group's rows: a sink for the sub-block update's out-of-range stores
    const size_t lds = sizeof(T) * (size_t)(rmax + 1) * (NB + 1);
    if (!attr_set[sizeof(T) == 8])
    {
        PVL_HIP(hipFuncSetAttribute((const void *)panel_kernel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
        attr_set[sizeof(T) == 8] = true;
    }
    PanelX<T> *X = (PanelX<T> *)g_ws.panelx;
    int *status = g_ws.err;
    factor_reset_kernel<<<1, 256, 0, stream()>>>(&X->rec, status);
    PVL_HIP(hipGetLastError());
    unsigned base = 0;
    // Panel width: 32 up to n = 2048, where single-workgroup register panels cover every
    // panel and the trailing updates are latency-bound anyway; 64 above.
    const int nbf = !pivoting ? NB : (n <= 2048 ? kPanelWgCols : NB);
    // Outer blocks of 4 panels for n > 2048: the trailing matrix right of and below an outer
    // block is updated by one GEMM covering K = 256 (a quarter of the C traffic of four
    // K = 64 updates; each entry subtracts one sum of the 256 products);
    // inside the block each panel updates the rest of the block row (all columns to the
    // right) and the block's remaining columns below it.  For n <= 2048 the outer block is
    // one panel and the updates reduce to the usual one per panel.
    const int nbo = n > 2048 ? 4 * nbf : nbf;
    auto gemm = [&](int row0, int m2, int col0, int n2, int kk0, int K) -> int {
        if (m2 <= 0 || n2 <= 0 || K <= 0)
            return PVL_OK;
        const int nt = (n2 + 63) / 64, cpx = (nt + kXcds - 1) / kXcds;
        const int mt = (m2 + 63) / 64;
        ProfScope prof(sizeof(T) == 8 ? "lu_gemm_f64" : "lu_gemm_f32", 2.0 * m2 * n2 * K);
        if (K > NB || (nbo > nbf && std::max(m2, n2) >= 2048))
            // an outer block's trailing update (one sum over the block's K columns), and for
            // n > 2048 the long updates inside an outer block
            return blas_update<T>(A + (size_t)row0 * n + col0, A + (size_t)row0 * n + kk0, A + (size_t)kk0 * n + col0, n,
                                  m2, n2, K);
        bool vec = false;
        if constexpr (sizeof(T) == 4)
            vec = n % 4 == 0;
        if (vec)
            gemm_kernel<float, true><<<kXcds * mt * cpx, 256, 0, stream()>>>((float *)A, n, row0, m2, col0, n2, kk0,
                                                                             K, cpx);
        else
            gemm_kernel<T, false><<<kXcds * mt * cpx, 256, 0, stream()>>>(A, n, row0, m2, col0, n2, kk0, K, cpx);
        PVL_HIP(hipGetLastError());
        return PVL_OK;
    };
    // U12 of the panel at k for columns [c0, c1)
    auto trsm = [&](int k, int c0, int c1) -> int {
        if (c1 <= c0)
            return PVL_OK;
        const int jb = std::min(nbf, n - k);
        ProfScope prof(sizeof(T) == 8 ? "lu_trsm_f64" : "lu_trsm_f32");
        if (jb == NB)
            trsm_kernel<T, NB><<<(c1 - c0 + 63) / 64, 256, 0, stream()>>>(A, n, c0, c1, k);
        else
            trsm_kernel<T, kPanelWgCols><<<(c1 - c0 + 63) / 64, 256, 0, stream()>>>(A, n, c0, c1, k);
        PVL_HIP(hipGetLastError());
        return PVL_OK;
    };
    for (int ko = 0; ko < n; ko += nbo)
    {
    const int kend = std::min(n, ko + nbo);
    for (int k = ko; k < kend; k += nbf)
    {
        const int jb = std::min(nbf, n - k);
        const int m = n - k;
        if (!pivoting)
        {
            T *P = A + (size_t)k * n + k;
            ProfScope prof(sizeof(T) == 8 ? "lu_panel_f64" : "lu_panel_f32");
            panel_nopiv_kernel<T><<<std::max(1, (m - jb + kNopivRows - 1) / kNopivRows), kNopivThreads, 0,
                                     stream()>>>(P, n, m, jb, k, piv, status);
            PVL_HIP(hipGetLastError());
        }
        else if (nbf == kPanelWgCols && m <= panel_wg_rows(sizeof(T)))
        {
            T *P = A + (size_t)k * n + k;
            ProfScope prof(sizeof(T) == 8 ? "lu_panel_f64" : "lu_panel_f32");
            if (int st = panel_wg(P, n, m, jb, k, piv, status))
                return st;
        }
        else
        {
            // rows per group: as many groups as needed for LDS, at most rmax rows each
            const int G = (m + rmax - 1) / rmax;
            const int R = (m + G - 1) / G;
            if (G > kMaxGroups)
            {
                set_error("matrix too tall for the panel kernel (%d rows)", m);
                return PVL_ERR_ARG;
            }
            T *P = A + (size_t)k * n + k;
            int lda = n, row_base = k;
            int mm = m, jbb = jb, RR = R;
            unsigned bb = base;
            void *args[] = {&P, &lda, &mm, &jbb, &RR, &row_base, &piv, &X, &bb, &status};
            base += (unsigned)jb; // one exchange round per column, consecutive over the launches
            const size_t lds_now = sizeof(T) * (size_t)(R + 1) * (NB + 1);
            {
                ProfScope prof(sizeof(T) == 8 ? "lu_panel_f64" : "lu_panel_f32");
                if (int st = launch_grid_synced((const void *)panel_kernel<T>, G, 256, args, lds_now, stream()))
                    return st;
            }
        }
        // info / err readback is deferred to the end (one sync per factorisation)
        if (pivoting)
        {
            ProfScope prof(sizeof(T) == 8 ? "lu_laswp_f64" : "lu_laswp_f32");
            laswp_kernel<T><<<(n + 255) / 256, 256, 0, stream()>>>(A, n, n, k, jb, piv);
            PVL_HIP(hipGetLastError());
        }
        if (pivoting)
        {
            // U12 and the update within the outer block's columns only: the block row right
            // of the outer block is formed after all of the block's interchanges (as blocked
            // dgetrf does), since a later panel's swap may bring up a row from below
            if (int st = trsm(k, k + jb, kend))
                return st;
            if (int st = gemm(k + jb, n - (k + jb), k + jb, kend - (k + jb), k, jb))
                return st;
        }
        else if (n - k - jb > 0)
        {
            // no interchanges: the panel kernel formed U12 for every column to the right, so
            // the rest of the block row is updated across all columns right away
            if (int st = gemm(k + jb, kend - (k + jb), k + jb, n - (k + jb), k, jb))
                return st;
            if (int st = gemm(kend, n - kend, k + jb, kend - (k + jb), k, jb))
                return st;
        }
    }
    if (pivoting)
    {
        // block row right of the outer block: U12 = L11^-1 A12 panel by panel
        for (int k = ko; k < kend; k += nbf)
        {
            const int jb = std::min(nbf, n - k);
            if (int st = trsm(k, kend, n))
                return st;
            if (int st = gemm(k + jb, kend - (k + jb), kend, n - kend, k, jb))
                return st;
        }
    }
    // the trailing matrix, once per outer block
    if (int st = gemm(kend, n - kend, kend, n - kend, ko, kend - ko))
        return st;
    }
    status_to_host_kernel<<<1, 64, 0, stream()>>>(status, g_status_host_dev);
    PVL_HIP(hipGetLastError());
    PVL_HIP(hipStreamSynchronize(stream()));
    if (int st = take_async_error())
        return st;
    int st2[4];
    for (int i = 0; i < 4; ++i)
        st2[i] = g_status_host[i];
    if (st2[1])
    {
        set_error("LU panel exchange timed out waiting for another workgroup's record");
        return PVL_ERR_HIP;
    }
    *info_out = pivoting ? st2[0] : (st2[2] == 0x7F7F7F7F ? 0 : st2[2]);
    if (would_pivot)
        *would_pivot = !pivoting && st2[3] != 0;
    return PVL_OK;
}

} // namespace

// Shared with the solver entry points below and other translation units.
struct pvl_lu
{
    int n = 0;
    int prec = PVL_FP64;
    void *lu = nullptr;    // factors (T), owned unless borrowed
    bool own_lu = false;
    int *piv = nullptr;    // device, n
    int *perm = nullptr;   // device, n: row i of P B is row perm[i] of B
    void *work = nullptr;  // n * kMaxRhs (T) staging for permuted right-hand sides
    int pivoting = PVL_PIVOT_PARTIAL;
    float *dinv = nullptr; // FP32: inverses of the 128 x 128 diagonal blocks of L and U
    bool dinv_valid = false;
    unsigned *flow = nullptr; // tickets and block flags of the single-launch solves
    unsigned epoch = 0;       // flag value of the latest solve launch
};

namespace
{

template <typename T> int finish_factor(pvl_lu *h, bool interchanged)
{
    const int n = h->n;
    if constexpr (sizeof(T) == 4)
    {
        // Only solves whose right-hand sides exceed one workgroup's LDS use the inverses.
        if ((size_t)n * kMaxRhs * sizeof(float) > (size_t)kTrsvWgBytes)
        {
            const int nblk = (n + kFlowBs32 - 1) / kFlowBs32;
            if (!h->dinv)
                PVL_HIP(hipMalloc(&h->dinv, sizeof(float) * 2 * (size_t)nblk * kFlowBs32 * kFlowBs32));
            ProfScope prof("lu_diag_inverse_f32");
            diag_inverse_kernel<kFlowBs32><<<dim3(nblk, 2), 256, 0, stream()>>>((const float *)h->lu, n, n, h->dinv);
            PVL_HIP(hipGetLastError());
            h->dinv_valid = true;
        }
    }
    if (!interchanged)
    {
        identity_kernel<<<(n + 255) / 256, 256, 0, stream()>>>(h->perm, n);
        PVL_HIP(hipGetLastError());
        return PVL_OK;
    }
    // permutation from the interchange list (LAPACK order), computed on the host
    std::vector<int> piv(n), perm(n);
    PVL_HIP(hipMemcpyAsync(piv.data(), h->piv, sizeof(int) * n, hipMemcpyDeviceToHost, stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    for (int i = 0; i < n; ++i)
        perm[i] = i;
    for (int i = 0; i < n; ++i)
        std::swap(perm[i], perm[piv[i]]);
    PVL_HIP(hipMemcpyAsync(h->perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    return PVL_OK;
}

template <typename T> int solve_impl(pvl_lu *h, T *B, int nrhs)
{
    const int n = h->n;
    if (int st = ensure_ws())
        return st;
    T *work = (T *)h->work;
    for (int q0 = 0; q0 < nrhs; q0 += kMaxRhs)
    {
        const int nq = std::min(kMaxRhs, nrhs - q0);
        // The solve kernels work on a contiguous n x nq block: gather (and permute) into work.
        if (nq == nrhs)
        {
            permute_rows_kernel<T><<<(n * nq + 255) / 256, 256, 0, stream()>>>(B, work, n, nq, h->perm);
        }
        else
        {
            set_error("more than %d right-hand sides per call are not supported", kMaxRhs);
            return PVL_ERR_ARG;
        }
        PVL_HIP(hipGetLastError());
        const size_t xbytes = sizeof(T) * (size_t)n * nq;
        if (nq == 1 && xbytes <= kTrsvLaMaxX)
        {
            ProfScope prof(sizeof(T) == 8 ? "lu_trsv_f64" : "lu_trsv_f32");
            const T *Ap = (const T *)h->lu;
            trsv_la_kernel<T, true, kTrsvLaBs><<<1, kTrsvLaThreads, xbytes, stream()>>>(Ap, n, n, work);
            trsv_la_kernel<T, false, kTrsvLaBs><<<1, kTrsvLaThreads, xbytes, stream()>>>(Ap, n, n, work);
            PVL_HIP(hipGetLastError());
        }
        else if (xbytes <= (size_t)kTrsvWgBytes)
        {
            ProfScope prof(sizeof(T) == 8 ? "lu_trsv_f64" : "lu_trsv_f32");
            const T *Ap = (const T *)h->lu;
            trsv_wg_kernel<T, true><<<1, kTrsvWgThreads, xbytes, stream()>>>(Ap, n, n, work, nq);
            trsv_wg_kernel<T, false><<<1, kTrsvWgThreads, xbytes, stream()>>>(Ap, n, n, work, nq);
            PVL_HIP(hipGetLastError());
        }
        else
        {
            if (sizeof(T) == 4 && !h->dinv_valid)
            {
                set_error("solve: the FP32 factorisation has no diagonal-block inverses");
                return PVL_ERR_ARG;
            }
            ProfScope prof(sizeof(T) == 8 ? "lu_trsv_f64" : "lu_trsv_f32");
            const T *Ap = (const T *)h->lu;
            flow_reset_kernel<<<1, 64, 0, stream()>>>(h->flow, g_ws.err + 5);
            // FP32: 128-row blocks (half the dependent steps of 64), 512 threads
            constexpr int BS = sizeof(T) == 4 ? kFlowBs32 : NB, NT = sizeof(T) == 4 ? 512 : 256;
            const int nb = (n + BS - 1) / BS;
            for (int pass = 0; pass < 2; ++pass)
            {
                if (++h->epoch == 0)
                    h->epoch = 1;
                if (pass == 0)
                    trsv_flow_kernel<T, true, BS, NT><<<nb, NT, 0, stream()>>>(Ap, n, n, work, nq, h->dinv, h->flow,
                                                                              h->epoch, g_ws.err + 5);
                else
                    trsv_flow_kernel<T, false, BS, NT><<<nb, NT, 0, stream()>>>(Ap, n, n, work, nq, h->dinv, h->flow,
                                                                               h->epoch, g_ws.err + 5);
            }
            record_async_error_kernel<<<1, 64, 0, stream()>>>(g_ws.err + 5, g_status_host_dev);
            PVL_HIP(hipGetLastError());
        }
        if (int e = device_copy(B, work, sizeof(T) * (size_t)n * nq))
            return e;
    }
    return PVL_OK;
}

int alloc_handle(pvl_lu *h, int n, int prec)
{
    const size_t es = prec == PVL_FP64 ? 8 : 4;
    PVL_HIP(hipMalloc(&h->piv, sizeof(int) * n));
    PVL_HIP(hipMalloc(&h->perm, sizeof(int) * n));
    PVL_HIP(hipMalloc(&h->work, es * (size_t)n * kMaxRhs));
    const int nblk = (n + NB - 1) / NB;
    PVL_HIP(hipMalloc(&h->flow, sizeof(unsigned) * (kFlowHeader + nblk)));
    return device_fill32(h->flow, 0u, sizeof(unsigned) * (kFlowHeader + nblk));
}

} // namespace

extern "C" {

int pvl_lu_create(pvl_lu **out, int n, int precision)
{
    PVL_API_LOCK;
    *out = nullptr;
    if (int st = ensure_init())
        return st;
    if (n <= 0 || (precision != PVL_FP64 && precision != PVL_FP32))
    {
        set_error("lu_create: bad n=%d or precision=%d", n, precision);
        return PVL_ERR_ARG;
    }
    if ((n + NB - 1) / NB > 2 * kMaxGroups * 2)
    {
        set_error("lu_create: n=%d exceeds the supported size", n);
        return PVL_ERR_ARG;
    }
    pvl_lu *h = new pvl_lu;
    h->n = n;
    h->prec = precision;
    const size_t es = precision == PVL_FP64 ? 8 : 4;
    int st = check_hip(hipMalloc(&h->lu, es * (size_t)n * n), "hipMalloc(lu)");
    h->own_lu = true;
    if (!st)
        st = alloc_handle(h, n, precision);
    if (st)
    {
        pvl_lu_destroy(h);
        return st;
    }
    *out = h;
    return PVL_OK;
}

int pvl_lu_destroy(pvl_lu *h)
{
    PVL_API_LOCK;
    if (!h)
        return PVL_OK;
    (void)hipStreamSynchronize(stream());
    if (h->own_lu)
        (void)hipFree(h->lu);
    (void)hipFree(h->piv);
    (void)hipFree(h->perm);
    (void)hipFree(h->work);
    (void)hipFree(h->dinv);
    (void)hipFree(h->flow);
    delete h;
    return PVL_OK;
}

// Partial pivoting (default; LAPACK's interchanges) or none: the unpivoted factorisation
// has no row search, so its panels need no grid-wide step per column; it is meant for an
// FP32 factorisation that FP64 refinement corrects, on systems whose diagonal dominates
// enough (refinement reports PVL_ERR_NOTCONV otherwise).
int pvl_lu_set_pivoting(pvl_lu *h, int pivoting)
{
    PVL_API_LOCK;
    if (!h || (pivoting != PVL_PIVOT_NONE && pivoting != PVL_PIVOT_PARTIAL && pivoting != PVL_PIVOT_AUTO))
    {
        set_error("lu_set_pivoting: %s", h ? "unknown mode" : "null LU handle");
        return PVL_ERR_ARG;
    }
    h->pivoting = pivoting;
    return PVL_OK;
}

// Factor: copies A (n x n row-major, FP64 on the device) into the handle's storage in the
// handle's precision and factors it.  info_out = first zero pivot (1-based) or 0.
int pvl_lu_factor(pvl_lu *h, const double *A_dev, int *info_out)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    const int n = h->n;
    const size_t cnt = (size_t)n * n;
    int info = 0, st;
    // PVL_PIVOT_AUTO: factor without interchanges; if partial pivoting would have chosen
    // another row at any column, factor again with partial pivoting.  Either way the
    // result is partial pivoting's factorisation.
    const bool try_free = h->pivoting != PVL_PIVOT_PARTIAL;
    bool pivoted = false;
    h->dinv_valid = false;
    for (int attempt = 0; attempt < 2; ++attempt)
    {
        const bool pivot = attempt == 1 || !try_free;
        pivoted = pivot;
        bool would_pivot = false;
        if (h->prec == PVL_FP64)
        {
            if (int e = device_copy(h->lu, A_dev, sizeof(double) * cnt))
                return e;
            st = factor_impl<double>(n, (double *)h->lu, h->piv, &info, pivot, &would_pivot);
        }
        else
        {
            f64_to_f32_kernel<<<2048, 256, 0, stream()>>>(A_dev, (float *)h->lu, cnt);
            PVL_HIP(hipGetLastError());
            st = factor_impl<float>(n, (float *)h->lu, h->piv, &info, pivot, &would_pivot);
        }
        if (st || pivot || !(h->pivoting == PVL_PIVOT_AUTO && would_pivot))
            break;
    }
    if (!st)
        st = h->prec == PVL_FP64 ? finish_factor<double>(h, pivoted) : finish_factor<float>(h, pivoted);
    if (info_out)
        *info_out = info;
    if (!st && info)
    {
        set_error("exactly singular: U(%d,%d) is zero", info - 1, info - 1);
        return PVL_ERR_SINGULAR;
    }
    return st;
}

// Solve A X = B in the handle's precision; B (n x nrhs row-major, same precision as the
// handle) is overwritten with X.
int pvl_lu_solve(pvl_lu *h, void *B_dev, int nrhs)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (nrhs <= 0 || nrhs > kMaxRhs)
    {
        set_error("lu_solve: nrhs=%d out of range 1..%d", nrhs, kMaxRhs);
        return PVL_ERR_ARG;
    }
    if (h->prec == PVL_FP64)
        return solve_impl<double>(h, (double *)B_dev, nrhs);
    return solve_impl<float>(h, (float *)B_dev, nrhs);
}

int pvl_lu_pivots(const pvl_lu *h, int *piv_host)
{
    PVL_API_LOCK;
    PVL_HIP(hipMemcpyAsync(piv_host, h->piv, sizeof(int) * h->n, hipMemcpyDeviceToHost, stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    return PVL_OK;
}

int pvl_lu_factors(const pvl_lu *h, void *lu_dev_out)
{
    PVL_API_LOCK;
    const size_t es = h->prec == PVL_FP64 ? 8 : 4;
    PVL_HIP(hipMemcpyAsync(lu_dev_out, h->lu, es * (size_t)h->n * h->n, hipMemcpyDeviceToDevice, stream()));
    return PVL_OK;
}

// R = B - A X in FP64 (A FP64, n x n; X, B, R n x nrhs).
int pvl_residual(int n, const double *A_dev, const double *X_dev, const double *B_dev, double *R_dev, int nrhs)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (nrhs <= 0 || nrhs > kMaxRhs)
    {
        set_error("residual: nrhs=%d out of range", nrhs);
        return PVL_ERR_ARG;
    }
    // K slices: enough waves to fill the device, at least 256 columns each
    const int row_tiles = (n + 15) / 16;
    int slices = (8192 + row_tiles - 1) / row_tiles;
    slices = std::max(1, std::min(slices, (n + 255) / 256));
    const int kslice = (((n + slices - 1) / slices) + 15) & ~15;
    slices = (n + kslice - 1) / kslice;
    const size_t count = (size_t)n * nrhs;
    double *part = (double *)scratch(5, sizeof(double) * count * slices);
    if (!part)
        return PVL_ERR_NOMEM;
    ProfScope prof("residual_f64");
    const dim3 grid((row_tiles + 3) / 4, slices);
    if (nrhs <= 16)
        residual_mfma_kernel<1><<<grid, 256, 0, stream()>>>(A_dev, n, n, X_dev, nrhs, kslice, part);
    else
        residual_mfma_kernel<2><<<grid, 256, 0, stream()>>>(A_dev, n, n, X_dev, nrhs, kslice, part);
    PVL_HIP(hipGetLastError());
    residual_sum_kernel<<<(unsigned)((count + 255) / 256), 256, 0, stream()>>>(part, slices, count, B_dev, R_dev);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

// Mixed precision: FP32 LU of A (handle h, precision FP32, already factored from A_dev),
// then FP64 iterative refinement X <- X + solve32(B - A X), starting from X = 0, until every
// column's FP64 residual satisfies max|r_q| <= tol * max|b_q|.  Stops early, reporting
// PVL_ERR_NOTCONV, when the worst relative residual fails to halve twice in a row (the
// FP32 factors are too inaccurate for this matrix) or is not finite.  iters_out:
// corrections applied; res_out[nrhs]: final max|r_q| / max|b_q| per column.
int pvl_solve_refine(pvl_lu *h, const double *A_dev, const double *B_dev, double *X_dev, int nrhs, double tol,
                     int max_iter, int *iters_out, double *res_out)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (h->prec != PVL_FP32)
    {
        set_error("solve_refine needs an FP32 factorisation");
        return PVL_ERR_ARG;
    }
    const int n = h->n;
    if (nrhs <= 0 || nrhs > kMaxRhs || max_iter < 1)
    {
        set_error("solve_refine: nrhs=%d out of range 1..%d (max_iter=%d)", nrhs, kMaxRhs, max_iter);
        return PVL_ERR_ARG;
    }
    const size_t cnt = (size_t)n * nrhs;
    double *R = (double *)scratch(3, sizeof(double) * cnt);
    float *D32 = (float *)scratch(4, sizeof(float) * cnt);
    double *norms = (double *)scratch(6, sizeof(double) * 2 * nrhs);
    if (!R || !D32 || !norms)
        return PVL_ERR_NOMEM;
    ProfScope prof_all("solve_refine");
    if (int e = device_fill32(X_dev, 0u, sizeof(double) * cnt))
        return e;
    // norms[nrhs + q] = max|b_q| (once); norms[q] = max|r_q| of the current iterate
    colabsmax_kernel<<<nrhs, 256, 0, stream()>>>(B_dev, n, nrhs, norms + nrhs);
    PVL_HIP(hipGetLastError());
    std::vector<double> hn(2 * nrhs);
    int it = 0, stalls = 0;
    double prev_worst = INFINITY, worst = INFINITY;
    bool converged = false;
    for (;;)
    {
        // X = 0 at the start, so the first residual is B itself
        const double *Rk = B_dev;
        if (it > 0)
        {
            if (int st = pvl_residual(n, A_dev, X_dev, B_dev, R, nrhs))
                return st;
            Rk = R;
        }
        colabsmax_kernel<<<nrhs, 256, 0, stream()>>>(Rk, n, nrhs, norms);
        PVL_HIP(hipGetLastError());
        PVL_HIP(hipMemcpyAsync(hn.data(), norms, sizeof(double) * 2 * nrhs, hipMemcpyDeviceToHost, stream()));
        PVL_HIP(hipStreamSynchronize(stream()));
        if (int st = take_async_error())
            return st;
        worst = 0.0;
        for (int q = 0; q < nrhs; ++q)
        {
            const double r = hn[nrhs + q] > 0 ? hn[q] / hn[nrhs + q] : hn[q];
            if (res_out)
                res_out[q] = r;
            worst = r <= DBL_MAX ? std::max(worst, r) : INFINITY;
        }
        if (worst <= tol)
        {
            converged = true;
            break;
        }
        stalls = worst > 0.5 * prev_worst ? stalls + 1 : 0;
        prev_worst = worst;
        if (it >= max_iter || stalls >= 2 || !(worst <= DBL_MAX))
            break;
        f64_to_f32_kernel<<<512, 256, 0, stream()>>>(Rk, D32, cnt);
        PVL_HIP(hipGetLastError());
        if (int st = solve_impl<float>(h, D32, nrhs))
            return st;
        add_f32_kernel<<<512, 256, 0, stream()>>>(X_dev, D32, cnt);
        PVL_HIP(hipGetLastError());
        ++it;
    }
    if (iters_out)
        *iters_out = it;
    if (!converged)
    {
        set_error("iterative refinement stopped at max|r|/max|b| = %.3e (target %.1e) after %d corrections", worst,
                  tol, it);
        return PVL_ERR_NOTCONV;
    }
    return PVL_OK;
}

int pvl_convert_f64_f32(const double *src, float *dst, size_t count)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (count == 0)
        return PVL_OK;
    f64_to_f32_kernel<<<1024, 256, 0, stream()>>>(src, dst, count);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

// Closed-body deflation: A[i][j] += c for i, j in [r0, r1) (rank-one shift c 1 1^T on the
// body's block; removes the uniform-circulation null direction of a closed surface).
__global__ void deflate_kernel(double *A, int n, int r0, int r1, double c)
{
    main_stream_priority();
    const int len = r1 - r0;
    for (size_t idx = blockIdx.x * (size_t)256 + threadIdx.x; idx < (size_t)len * len;
         idx += (size_t)gridDim.x * 256)
    {
        const int i = r0 + (int)(idx / len), j = r0 + (int)(idx % len);
        A[(size_t)i * n + j] += c;
    }
}

int pvl_deflate(int n, double *A_dev, int r0, int r1, double c)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (r0 < 0 || r1 > n || r0 >= r1)
    {
        set_error("deflate: bad range [%d, %d) for n=%d", r0, r1, n);
        return PVL_ERR_ARG;
    }
    deflate_kernel<<<2048, 256, 0, stream()>>>(A_dev, n, r0, r1, c);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

} // extern "C"
