// FP32 restarted GMRES inside FP64 iterative refinement, for the circulation system
// A x = b of one step.  No reference counterpart: the reference factors A with LAPACK every
// step (python/pyvl/solver.py:223-227); here a prediction of x (e.g. the learned one)
// warm-starts the refinement, and a solve that does not converge reports so the caller can
// fall back to LU.
//
// One GMRES(m) cycle is a single grid-wide kernel: workgroup g owns a block of rows;
// per Arnoldi step it forms w = A M^-1 v_j for its rows (M = diag(A), right Jacobi
// preconditioning), orthogonalises with classical Gram-Schmidt applied twice (partial dot
// products of all groups combined after a grid barrier), normalises, and every group
// applies the same Givens rotations to its private copy of the Hessenberg column, so the
// residual estimate and the stopping decision are identical everywhere without a host
// round trip.  Four grid barriers per step (two Gram-Schmidt passes, the norm, and the new
// basis vector).
#include "pvl_internal.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

using namespace pvl;

namespace
{

constexpr int kMaxM = 48;          // restart length cap
constexpr int kGroups = 128;       // grid size cap (every group resident at once)
constexpr unsigned kSpin = 20000000u;

__device__ __forceinline__ void gbar(unsigned *counter, unsigned target, int *err)
{
    // every wave's stores complete; thread 0's release (an L2 write-back) covers the workgroup
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(counter, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target)
        {
            __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpin || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
}

__device__ __forceinline__ float ld(const float *p)
{
    return __builtin_bit_cast(float, __hip_atomic_load((const unsigned *)p, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT));
}
__device__ __forceinline__ void st(float *p, float v)
{
    __hip_atomic_store((unsigned *)p, __builtin_bit_cast(unsigned, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct GmresX
{
    unsigned counter;
    int err;
    int pad[2];
    float part[3][kGroups][kMaxM + 2]; // per-group partial sums, one buffer per barrier phase
};

// rows of group g: [g*rpg, min(n, (g+1)*rpg)); thread t handles rows r0 + t, r0 + t + 256, ...
__global__ void __launch_bounds__(256) gmres_cycle_kernel(int n, const float *__restrict__ A,
                                                          const float *__restrict__ dinv,
                                                          const float *__restrict__ rhs, float *__restrict__ d,
                                                          float *__restrict__ V, int m, float tol, int rpg,
                                                          GmresX *__restrict__ X, int *__restrict__ result)
{
    __shared__ float Hs[kMaxM + 1][kMaxM];    // Hessenberg (this group's copy)
    __shared__ float cs[kMaxM], sn[kMaxM], gv[kMaxM + 1];
    __shared__ float red[kMaxM + 8];
    __shared__ float hcol[kMaxM + 2];
    __shared__ int s_steps;
    __shared__ float s_beta;
    const int tid = threadIdx.x, g = blockIdx.x, G = gridDim.x;
    const int r0 = g * rpg, r1 = min(n, r0 + rpg);
    unsigned bar = 0;
    // local rows' w in registers: at most rpg / 256 rows per thread (rpg <= 256)
    // 1. beta = ||rhs||, v0 = rhs / beta
    float s = 0.0f;
    for (int r = r0 + tid; r < r1; r += 256)
        s += rhs[r] * rhs[r];
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off);
    if ((tid & 63) == 0)
        red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0)
        st(&X->part[0][g][0], red[0] + red[1] + red[2] + red[3]);
    bar += G;
    gbar(&X->counter, bar, &X->err);
    if (tid == 0)
    {
        float tot = 0.0f;
        for (int q = 0; q < G; ++q)
            tot += ld(&X->part[0][q][0]);
        s_beta = sqrtf(tot);
        gv[0] = s_beta;
    }
    __syncthreads();
    const float beta = s_beta;
    if (!(beta > 0.0f))
    {
        for (int r = r0 + tid; r < r1; r += 256)
            d[r] = 0.0f;
        if (g == 0 && tid == 0)
            result[0] = 0;
        return;
    }
    for (int r = r0 + tid; r < r1; r += 256)
        st(&V[r], rhs[r] / beta);
    bar += G;
    gbar(&X->counter, bar, &X->err);
    int j = 0;
    float resid = beta;
    for (; j < m; ++j)
    {
        const float *vj = V + (size_t)j * n;
        float *vn = V + (size_t)(j + 1) * n;
        // w = A (dinv .* v_j) for own rows; one wave per row, lanes over columns
        const int lane = tid & 63, wv = tid >> 6;
        for (int r = r0 + wv; r < r1; r += 4)
        {
            const float *a = A + (size_t)r * n;
            float acc = 0.0f;
            for (int c = lane; c < n; c += 64)
                acc = fmaf(a[c], dinv[c] * ld(&vj[c]), acc);
            for (int off = 32; off > 0; off >>= 1)
                acc += __shfl_xor(acc, off);
            if (lane == 0)
                vn[r] = acc; // staged in V[j+1] (own rows only)
        }
        __syncthreads();
        // two passes of classical Gram-Schmidt: h_i = <w, v_i> summed over groups, w -= h_i v_i
        for (int pass = 0; pass < 2; ++pass)
        {
            for (int i = wv; i <= j; i += 4)
            {
                float acc = 0.0f;
                const float *vi = V + (size_t)i * n;
                for (int r = r0 + lane; r < r1; r += 64)
                    acc = fmaf(vn[r], ld(&vi[r]), acc);
                for (int off = 32; off > 0; off >>= 1)
                    acc += __shfl_xor(acc, off);
                if (lane == 0)
                    st(&X->part[pass][g][i], acc);
            }
            bar += G;
            gbar(&X->counter, bar, &X->err);
            for (int i = tid; i <= j; i += 256)
            {
                float tot = 0.0f;
                for (int q = 0; q < G; ++q)
                    tot += ld(&X->part[pass][q][i]);
                red[i] = tot;
                hcol[i] = pass == 0 ? tot : hcol[i] + tot;
            }
            __syncthreads();
            for (int r = r0 + tid; r < r1; r += 256)
            {
                float wr = vn[r];
                for (int i = 0; i <= j; ++i)
                    wr = fmaf(-red[i], ld(&V[(size_t)i * n + r]), wr);
                vn[r] = wr;
            }
            __syncthreads();
        }
        // ||w||
        {
            float acc = 0.0f;
            for (int r = r0 + tid; r < r1; r += 256)
                acc = fmaf(vn[r], vn[r], acc);
            for (int off = 32; off > 0; off >>= 1)
                acc += __shfl_xor(acc, off);
            if ((tid & 63) == 0)
                red[kMaxM + 1 + (tid >> 6)] = acc;
            __syncthreads();
            if (tid == 0)
                st(&X->part[2][g][0], red[kMaxM + 1] + red[kMaxM + 2] + red[kMaxM + 3] + red[kMaxM + 4]);
            bar += G;
            gbar(&X->counter, bar, &X->err);
            if (tid == 0)
            {
                float tot = 0.0f;
                for (int q = 0; q < G; ++q)
                    tot += ld(&X->part[2][q][0]);
                hcol[j + 1] = sqrtf(tot);
            }
            __syncthreads();
        }
        const float hn = hcol[j + 1];
        for (int r = r0 + tid; r < r1; r += 256)
            st(&vn[r], hn > 0.0f ? vn[r] / hn : 0.0f);
        // Givens on this group's copy
        if (tid == 0)
        {
            for (int i = 0; i <= j + 1; ++i)
                Hs[i][j] = hcol[i];
            for (int i = 0; i < j; ++i)
            {
                const float a = Hs[i][j], b = Hs[i + 1][j];
                Hs[i][j] = cs[i] * a + sn[i] * b;
                Hs[i + 1][j] = -sn[i] * a + cs[i] * b;
            }
            const float a = Hs[j][j], b = Hs[j + 1][j];
            const float rr = sqrtf(a * a + b * b);
            cs[j] = rr > 0.0f ? a / rr : 1.0f;
            sn[j] = rr > 0.0f ? b / rr : 0.0f;
            Hs[j][j] = rr;
            Hs[j + 1][j] = 0.0f;
            gv[j + 1] = -sn[j] * gv[j];
            gv[j] = cs[j] * gv[j];
            s_steps = fabsf(gv[j + 1]) <= tol * beta || hn == 0.0f ? 1 : 0;
        }
        bar += G;
        gbar(&X->counter, bar, &X->err);
        resid = fabsf(gv[j + 1]);
        if (s_steps)
        {
            ++j;
            break;
        }
    }
    if (j > m)
        j = m;
    // y = H^-1 g (upper triangular j x j), then d = dinv .* (V y) for own rows
    __shared__ float y[kMaxM];
    if (tid == 0)
    {
        for (int i = j - 1; i >= 0; --i)
        {
            float t = gv[i];
            for (int k = i + 1; k < j; ++k)
                t -= Hs[i][k] * y[k];
            y[i] = t / Hs[i][i];
        }
    }
    __syncthreads();
    for (int r = r0 + tid; r < r1; r += 256)
    {
        float acc = 0.0f;
        for (int i = 0; i < j; ++i)
            acc = fmaf(y[i], ld(&V[(size_t)i * n + r]), acc);
        d[r] = dinv[r] * acc;
    }
    if (g == 0 && tid == 0)
    {
        result[0] = j;
        result[1] = __float_as_int(resid / beta);
    }
}

__global__ void diag_inv_kernel(const float *__restrict__ A, int n, float *__restrict__ dinv)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
    {
        const float a = A[(size_t)i * n + i];
        dinv[i] = a != 0.0f ? 1.0f / a : 1.0f;
    }
}

__global__ void r64_to_32_kernel(const double *__restrict__ r, float *__restrict__ r32, int n, double scale)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        r32[i] = (float)(r[i] * scale);
}

__global__ void add_scaled_kernel(double *__restrict__ x, const float *__restrict__ d, int n, double scale)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        x[i] += (double)d[i] * scale;
}

__global__ void maxabs_kernel(const double *__restrict__ r, int n, double *__restrict__ out)
{
    __shared__ double p[256];
    double m = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
        m = fabs(r[i]) <= DBL_MAX ? fmax(m, fabs(r[i])) : INFINITY; // NaN or Inf -> Inf
    p[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            p[threadIdx.x] = fmax(p[threadIdx.x], p[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMax((unsigned long long *)out, (unsigned long long)__double_as_longlong(p[0]));
}

struct GmresWs
{
    GmresX *x = nullptr;
    int *result = nullptr;
    int n_cap = 0, m_cap = 0;
    float *V = nullptr, *r32 = nullptr, *d32 = nullptr, *dinv = nullptr;
    double *r64 = nullptr, *norms = nullptr;
};
GmresWs g_gw;

} // namespace

void pvl::gmres_release_workspace()
{
    for (void *p : {(void *)g_gw.x, (void *)g_gw.result, (void *)g_gw.V, (void *)g_gw.r32, (void *)g_gw.d32,
                    (void *)g_gw.dinv, (void *)g_gw.r64, (void *)g_gw.norms})
        (void)hipFree(p);
    g_gw = GmresWs{};
}

namespace
{

int ensure_gmres_ws(int n, int m)
{
    if (!g_gw.x)
    {
        PVL_HIP(hipMalloc(&g_gw.x, sizeof(GmresX)));
        PVL_HIP(hipMalloc(&g_gw.result, 4 * sizeof(int)));
        PVL_HIP(hipMalloc(&g_gw.norms, 4 * sizeof(double)));
    }
    if (n > g_gw.n_cap || m > g_gw.m_cap)
    {
        n = std::max(n, g_gw.n_cap);
        m = std::max(m, g_gw.m_cap);
        (void)hipStreamSynchronize(stream());
        (void)hipFree(g_gw.V);
        (void)hipFree(g_gw.r32);
        (void)hipFree(g_gw.d32);
        (void)hipFree(g_gw.dinv);
        (void)hipFree(g_gw.r64);
        PVL_HIP(hipMalloc(&g_gw.V, sizeof(float) * (size_t)(m + 1) * n));
        PVL_HIP(hipMalloc(&g_gw.r32, sizeof(float) * n));
        PVL_HIP(hipMalloc(&g_gw.d32, sizeof(float) * n));
        PVL_HIP(hipMalloc(&g_gw.dinv, sizeof(float) * n));
        PVL_HIP(hipMalloc(&g_gw.r64, sizeof(double) * n));
        g_gw.n_cap = n;
        g_gw.m_cap = m;
    }
    return PVL_OK;
}

} // namespace

extern "C" {

// x (FP64, n) holds the initial guess and receives the solution of A x = b.  A64 (FP64)
// forms the residuals, A32 (FP32 copy of A) the inner GMRES(m) cycles.  Each outer step
// computes r = b - A x in FP64 and stops when max|r| <= tol * max|b|; otherwise one GMRES(m)
// cycle solves A d = r in FP32 (Jacobi right preconditioning) down to the relative residual
// max(inner_tol, tol max|b| / (2 max|r|)) -- no further than the final target needs -- and
// x += d.  At most max_outer corrections.  stats[0] = corrections, stats[1] = GMRES steps in
// total, stats[2] = 1 if converged; res_out[0] = max|b - A x0| / max|b| of the start,
// res_out[1] = the final one.  Returns PVL_ERR_NOTCONV when not converged (x then holds the
// last iterate) and PVL_OK for b = 0 (x = 0).
int pvl_gmres_refine(int n, const double *A64_dev, const float *A32_dev, const double *b_dev, double *x_dev, double tol,
                     int m, double inner_tol, int max_outer, int *stats, double *res_out)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (n <= 0 || m <= 0 || m > kMaxM || max_outer < 0)
    {
        set_error("gmres_refine: n=%d, m=%d (1..%d), max_outer=%d", n, m, kMaxM, max_outer);
        return PVL_ERR_ARG;
    }
    if (int st = ensure_gmres_ws(n, m))
        return st;
    const int G = std::min(kGroups, (n + 15) / 16);
    const int rpg = (n + G - 1) / G;
    if (rpg > 256 * 64)
    {
        set_error("gmres_refine: n=%d too large", n);
        return PVL_ERR_ARG;
    }
    ProfScope prof_all("gmres_refine");
    diag_inv_kernel<<<(n + 255) / 256, 256, 0, stream()>>>(A32_dev, n, g_gw.dinv);
    PVL_HIP(hipGetLastError());
    // norms[0] = max|b|, norms[2] = max|r| of the current outer step
    if (int e = device_fill32(g_gw.norms, 0u, 4 * sizeof(double)))
        return e;
    maxabs_kernel<<<64, 256, 0, stream()>>>(b_dev, n, g_gw.norms);
    PVL_HIP(hipGetLastError());
    int total_steps = 0, outer = 0, pending_cycle = 0;
    bool converged = false;
    double bmax = 0.0, rel0 = INFINITY, rel = INFINITY;
    for (;; ++outer)
    {
        if (int st = pvl_residual(n, A64_dev, x_dev, b_dev, g_gw.r64, 1))
            return st;
        if (int e = device_fill32(g_gw.norms + 2, 0u, sizeof(double)))
            return e;
        maxabs_kernel<<<64, 256, 0, stream()>>>(g_gw.r64, n, g_gw.norms + 2);
        PVL_HIP(hipGetLastError());
        // one host synchronisation per outer step: residual norm and the last cycle's result
        double hn[3];
        int hres[2] = {0, 0}, herr = 0;
        PVL_HIP(hipMemcpyAsync(hn, g_gw.norms, 3 * sizeof(double), hipMemcpyDeviceToHost, stream()));
        if (pending_cycle)
        {
            PVL_HIP(hipMemcpyAsync(hres, g_gw.result, 2 * sizeof(int), hipMemcpyDeviceToHost, stream()));
            PVL_HIP(hipMemcpyAsync(&herr, &g_gw.x->err, sizeof(int), hipMemcpyDeviceToHost, stream()));
        }
        PVL_HIP(hipStreamSynchronize(stream()));
        if (herr)
        {
            set_error("GMRES grid barrier timed out");
            return PVL_ERR_HIP;
        }
        total_steps += hres[0];
        pending_cycle = 0;
        bmax = hn[0];
        const double rmax = hn[2];
        if (bmax == 0.0)
        {
            // b = 0: the solution is x = 0
            if (int e = device_fill32(x_dev, 0u, sizeof(double) * n))
                return e;
            rel0 = rel = 0.0;
            converged = true;
            break;
        }
        rel = rmax / bmax;
        if (outer == 0)
            rel0 = rel;
        if (rel <= tol)
        {
            converged = true;
            break;
        }
        if (!(rmax <= DBL_MAX) || outer >= max_outer)
            break; // non-finite start or residual, or out of corrections: the caller falls back
        const double scale = 1.0 / rmax;
        r64_to_32_kernel<<<(n + 255) / 256, 256, 0, stream()>>>(g_gw.r64, g_gw.r32, n, scale);
        PVL_HIP(hipGetLastError());
        PVL_HIP(hipMemsetAsync(g_gw.x, 0, 16, stream()));
        int nn = n, mm = m, rp = rpg;
        float tl = (float)std::max(inner_tol, 0.5 * tol / rel);
        const float *Ap = A32_dev, *dv = g_gw.dinv, *rh = g_gw.r32;
        float *dd = g_gw.d32, *Vp = g_gw.V;
        GmresX *Xp = g_gw.x;
        int *res = g_gw.result;
        void *args[] = {&nn, (void *)&Ap, (void *)&dv, (void *)&rh, &dd, &Vp, &mm, &tl, &rp, &Xp, &res};
        {
            ProfScope prof("gmres_cycle");
            if (int st = launch_grid_synced((const void *)gmres_cycle_kernel, G, 256, args, 0, stream()))
                return st;
        }
        add_scaled_kernel<<<64, 256, 0, stream()>>>(x_dev, g_gw.d32, n, 1.0 / scale);
        PVL_HIP(hipGetLastError());
        pending_cycle = 1;
    }
    if (stats)
    {
        stats[0] = outer;
        stats[1] = total_steps;
        stats[2] = converged ? 1 : 0;
    }
    if (res_out)
    {
        res_out[0] = rel0;
        res_out[1] = rel;
    }
    if (!converged)
    {
        set_error("GMRES refinement did not reach %.1e (max|r|/max|b| = %.3e after %d corrections)", tol, rel, outer);
        return PVL_ERR_NOTCONV;
    }
    return PVL_OK;
}

} // extern "C"
