// Small MLP (BF16 operands, FP32 accumulation) trained and evaluated on the GPU with
// hand-written kernels: it predicts each panel's circulation to warm-start the iterative
// solve.  No reference counterpart (the reference solves every step with LAPACK).
//
// Network: in -> W1 -> SiLU -> W2 -> SiLU -> ... -> W_L -> out (scalar), width 256.
// Parameters are kept in FP32 (master copy, Adam moments) and mirrored in BF16 for the
// products.  Activations and back-propagated errors are stored in BF16; every product
// accumulates in FP32 on v_mfma_f32_16x16x32_bf16.
#include "pvl_internal.hpp"

#include <cmath>
#include <vector>

using namespace pvl;

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct pvl_mlp
{
    int d_in = 0, width = 0, n_layers = 0;
    std::vector<int> dims;          // n_layers + 1 entries
    std::vector<size_t> w_off, b_off; // offsets into the flat parameter vector
    size_t n_params = 0;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr; // FP32 master, grad, Adam
    bf16 *d_pb = nullptr;           // BF16 mirror
    int cap_batch = 0;
    std::vector<bf16 *> d_act;      // [n_layers] layer inputs A_l (batch x dims[l]) in BF16
    std::vector<float *> d_z;       // [n_layers] pre-activations Z_l (batch x dims[l+1]) FP32
    bf16 *d_delta[2] = {nullptr, nullptr}; // back-propagated errors (batch x width)
    float *d_out = nullptr;         // batch
    float *d_loss = nullptr;        // 1
    long step = 0;
};

namespace
{

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// C[m][j] = sum_r A(m, r) * B(j, r) with A(m, r) = A[m*am + r*ar], B(j, r) = B[j*bj + r*br]
// (BF16 operands, FP32 accumulation on MFMA 16x16x32).  64x64 tile per workgroup, 32x32
// per wave; K staged through LDS in steps of 32.  Epilogues:
//   EPI_FWD:  Z = C + bias[j] (FP32 out), Anext = SiLU(Z) in BF16 (if act) or Z
//   EPI_BWD:  D = C * SiLU'(Zprev[m][j])  -> BF16
//   EPI_WGRAD: G[m][j] = C (FP32)
enum
{
    EPI_FWD = 0,
    EPI_BWD = 1,
    EPI_WGRAD = 2
};

template <int EPI>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int R, const bf16 *__restrict__ A, long am,
                                                        long ar, const bf16 *__restrict__ B, long bj, long br,
                                                        const float *__restrict__ bias, float *__restrict__ outf,
                                                        bf16 *__restrict__ outb, const float *__restrict__ zprev,
                                                        int act)
{
    constexpr int KS = 32, LD = KS + 8;
    __shared__ bf16 As[64][LD];
    __shared__ bf16 Bs[64][LD];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    const int wm = w >> 1, wn = w & 1, lr = lane & 15, lk = lane >> 4;
    const int m0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
    f32x4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            acc[a][b] = f32x4{0, 0, 0, 0};
    for (int r0 = 0; r0 < R; r0 += KS)
    {
        for (int idx = tid; idx < 64 * KS; idx += 256)
        {
            const int i = idx / KS, k = idx % KS;
            const int m = m0 + i, j = j0 + i, r = r0 + k;
            As[i][k] = (m < M && r < R) ? A[m * am + r * ar] : (bf16)0.0f;
            Bs[i][k] = (j < N && r < R) ? B[j * bj + r * br] : (bf16)0.0f;
        }
        __syncthreads();
        bf16x8 af[2], bfr[2];
        for (int a = 0; a < 2; ++a)
            for (int e = 0; e < 8; ++e)
                af[a][e] = As[32 * wm + 16 * a + lr][8 * lk + e];
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 8; ++e)
                bfr[b][e] = Bs[32 * wn + 16 * b + lr][8 * lk + e];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b)
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        __syncthreads();
    }
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int v = 0; v < 4; ++v)
            {
                const int m = m0 + 32 * wm + 16 * a + 4 * lk + v;
                const int j = j0 + 32 * wn + 16 * b + lr;
                if (m >= M || j >= N)
                    continue;
                const float c = acc[a][b][v];
                const long o = (long)m * N + j;
                if constexpr (EPI == EPI_FWD)
                {
                    const float z = c + bias[j];
                    if (outf)
                        outf[o] = z;
                    if (outb)
                        outb[o] = (bf16)(act ? silu(z) : z);
                }
                else if constexpr (EPI == EPI_BWD)
                {
                    outb[o] = (bf16)(c * silu_grad(zprev[o]));
                }
                else
                {
                    outf[o] = c;
                }
            }
}

__global__ void f32_to_bf16_kernel(const float *__restrict__ a, bf16 *__restrict__ b, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
        b[i] = (bf16)a[i];
}

// Output layer error: d = 2 (y - t) / B (MSE), loss partials.
__global__ void __launch_bounds__(256) mse_kernel(const float *__restrict__ y, const float *__restrict__ t, int n,
                                                  bf16 *__restrict__ delta, float *__restrict__ loss)
{
    __shared__ float part[256];
    float s = 0.0f;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    {
        const float e = y[i] - t[i];
        delta[i] = (bf16)(2.0f * e / (float)n);
        s += e * e;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicAdd(loss, part[0] / (float)n);
}

// db[j] = sum_m delta[m][j]  (delta BF16, M x N)
__global__ void __launch_bounds__(256) bias_grad_kernel(const bf16 *__restrict__ delta, int M, int N,
                                                        float *__restrict__ db)
{
    const int j = blockIdx.x;
    __shared__ float part[256];
    float s = 0.0f;
    for (int m = threadIdx.x; m < M; m += 256)
        s += (float)delta[(long)m * N + j];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        db[j] = part[0];
}

__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, bf16 *__restrict__ pb, long n, float lr, float b1, float b2,
                            float eps, float c1, float c2, float wd)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256)
    {
        const float gi = g[i] + wd * p[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        p[i] = pi;
        pb[i] = (bf16)pi;
    }
}

int ensure_batch(pvl_mlp *h, int batch)
{
    if (batch <= h->cap_batch)
        return PVL_OK;
    (void)hipStreamSynchronize(stream());
    for (auto p : h->d_act)
        (void)hipFree(p);
    for (auto p : h->d_z)
        (void)hipFree(p);
    (void)hipFree(h->d_delta[0]);
    (void)hipFree(h->d_delta[1]);
    (void)hipFree(h->d_out);
    const int L = h->n_layers;
    h->d_act.assign(L, nullptr);
    h->d_z.assign(L, nullptr);
    for (int l = 0; l < L; ++l)
    {
        PVL_HIP(hipMalloc(&h->d_act[l], sizeof(bf16) * (size_t)batch * h->dims[l]));
        PVL_HIP(hipMalloc(&h->d_z[l], sizeof(float) * (size_t)batch * h->dims[l + 1]));
    }
    PVL_HIP(hipMalloc(&h->d_delta[0], sizeof(bf16) * (size_t)batch * h->width));
    PVL_HIP(hipMalloc(&h->d_delta[1], sizeof(bf16) * (size_t)batch * h->width));
    PVL_HIP(hipMalloc(&h->d_out, sizeof(float) * (size_t)batch));
    h->cap_batch = batch;
    return PVL_OK;
}

template <int EPI>
int launch_gemm(int M, int N, int R, const bf16 *A, long am, long ar, const bf16 *B, long bj, long br,
                const float *bias, float *outf, bf16 *outb, const float *zprev, int act)
{
    const dim3 grid((N + 63) / 64, (M + 63) / 64);
    gemm_bf16_kernel<EPI><<<grid, 256, 0, stream()>>>(M, N, R, A, am, ar, B, bj, br, bias, outf, outb, zprev, act);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

// Forward pass storing every layer's input (BF16) and pre-activation (FP32).
int forward(pvl_mlp *h, const float *X, int batch)
{
    if (int st = ensure_batch(h, batch))
        return st;
    const long nin = (long)batch * h->d_in;
    f32_to_bf16_kernel<<<256, 256, 0, stream()>>>(X, h->d_act[0], nin);
    PVL_HIP(hipGetLastError());
    const int L = h->n_layers;
    for (int l = 0; l < L; ++l)
    {
        const int K = h->dims[l], N = h->dims[l + 1];
        const bool last = l == L - 1;
        bf16 *next = last ? nullptr : h->d_act[l + 1];
        if (int st = launch_gemm<EPI_FWD>(batch, N, K, h->d_act[l], K, 1, h->d_pb + h->w_off[l], K, 1,
                                          h->d_p + h->b_off[l], h->d_z[l], next, nullptr, last ? 0 : 1))
            return st;
    }
    return PVL_OK;
}

// Backward pass of the MSE loss against T; gradients into d_g (FP32).
int backward(pvl_mlp *h, const float *T, int batch)
{
    const int L = h->n_layers;
    PVL_HIP(hipMemsetAsync(h->d_loss, 0, sizeof(float), stream()));
    // the last layer's pre-activation is the output (dims[L] == 1)
    mse_kernel<<<64, 256, 0, stream()>>>(h->d_z[L - 1], T, batch, h->d_delta[0], h->d_loss);
    PVL_HIP(hipGetLastError());
    int cur = 0;
    for (int l = L - 1; l >= 0; --l)
    {
        const int K = h->dims[l], N = h->dims[l + 1];
        bf16 *delta = h->d_delta[cur];
        // dW_l[n][k] = sum_b delta[b][n] * A_l[b][k]
        if (int st = launch_gemm<EPI_WGRAD>(N, K, batch, delta, 1, N, h->d_act[l], 1, K, nullptr,
                                            h->d_g + h->w_off[l], nullptr, nullptr, 0))
            return st;
        bias_grad_kernel<<<N, 256, 0, stream()>>>(delta, batch, N, h->d_g + h->b_off[l]);
        PVL_HIP(hipGetLastError());
        if (l > 0)
        {
            // delta_prev[b][k] = (sum_n delta[b][n] W_l[n][k]) * SiLU'(Z_{l-1}[b][k])
            if (int st = launch_gemm<EPI_BWD>(batch, K, N, delta, N, 1, h->d_pb + h->w_off[l], 1, K, nullptr,
                                              nullptr, h->d_delta[1 - cur], h->d_z[l - 1], 0))
                return st;
            cur = 1 - cur;
        }
    }
    return PVL_OK;
}

} // namespace

extern "C" {

int pvl_mlp_destroy(pvl_mlp *h);

int pvl_mlp_create(pvl_mlp **out, int d_in, int width, int n_layers)
{
    PVL_API_LOCK;
    *out = nullptr;
    if (int st = ensure_init())
        return st;
    if (d_in <= 0 || width <= 0 || n_layers < 2)
    {
        set_error("mlp_create: bad sizes d_in=%d width=%d layers=%d", d_in, width, n_layers);
        return PVL_ERR_ARG;
    }
    pvl_mlp *h = new pvl_mlp;
    h->d_in = d_in;
    h->width = width;
    h->n_layers = n_layers;
    h->dims.push_back(d_in);
    for (int l = 0; l < n_layers - 1; ++l)
        h->dims.push_back(width);
    h->dims.push_back(1);
    size_t off = 0;
    for (int l = 0; l < n_layers; ++l)
    {
        h->w_off.push_back(off);
        off += (size_t)h->dims[l] * h->dims[l + 1];
        h->b_off.push_back(off);
        off += (size_t)h->dims[l + 1];
    }
    h->n_params = off;
    int st = PVL_OK;
    auto alloc = [&](void **p, size_t bytes) {
        if (!st)
            st = check_hip(hipMalloc(p, bytes), "hipMalloc(mlp)");
    };
    alloc((void **)&h->d_p, sizeof(float) * off);
    alloc((void **)&h->d_g, sizeof(float) * off);
    alloc((void **)&h->d_m, sizeof(float) * off);
    alloc((void **)&h->d_v, sizeof(float) * off);
    alloc((void **)&h->d_pb, sizeof(bf16) * off);
    alloc((void **)&h->d_loss, sizeof(float));
    if (!st)
        st = check_hip(hipMemset(h->d_m, 0, sizeof(float) * off), "hipMemset");
    if (!st)
        st = check_hip(hipMemset(h->d_v, 0, sizeof(float) * off), "hipMemset");
    if (st)
    {
        pvl_mlp_destroy(h);
        return st;
    }
    *out = h;
    return PVL_OK;
}

int pvl_mlp_destroy(pvl_mlp *h)
{
    PVL_API_LOCK;
    if (!h)
        return PVL_OK;
    (void)hipStreamSynchronize(stream());
    (void)hipFree(h->d_p);
    (void)hipFree(h->d_g);
    (void)hipFree(h->d_m);
    (void)hipFree(h->d_v);
    (void)hipFree(h->d_pb);
    (void)hipFree(h->d_loss);
    for (auto p : h->d_act)
        (void)hipFree(p);
    for (auto p : h->d_z)
        (void)hipFree(p);
    (void)hipFree(h->d_delta[0]);
    (void)hipFree(h->d_delta[1]);
    (void)hipFree(h->d_out);
    delete h;
    return PVL_OK;
}

int pvl_mlp_num_params(const pvl_mlp *h, long *n)
{
    PVL_API_LOCK;
    *n = (long)h->n_params;
    return PVL_OK;
}

// Parameters (host FP32, layout: for each layer W (out x in, row-major) then b (out)).
int pvl_mlp_set_params(pvl_mlp *h, const float *params_host)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    PVL_HIP(hipMemcpyAsync(h->d_p, params_host, sizeof(float) * h->n_params, hipMemcpyHostToDevice, stream()));
    f32_to_bf16_kernel<<<256, 256, 0, stream()>>>(h->d_p, h->d_pb, (long)h->n_params);
    PVL_HIP(hipGetLastError());
    PVL_HIP(hipMemsetAsync(h->d_m, 0, sizeof(float) * h->n_params, stream()));
    PVL_HIP(hipMemsetAsync(h->d_v, 0, sizeof(float) * h->n_params, stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    h->step = 0;
    return PVL_OK;
}

int pvl_mlp_get_params(const pvl_mlp *h, float *params_host)
{
    PVL_API_LOCK;
    PVL_HIP(hipMemcpyAsync(params_host, h->d_p, sizeof(float) * h->n_params, hipMemcpyDeviceToHost, stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    return PVL_OK;
}

// Inference: Y[b] = net(X[b, :]) for X (batch x d_in, FP32 device), Y FP32 device.
int pvl_mlp_forward(pvl_mlp *h, const float *X_dev, float *Y_dev, int batch)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (batch <= 0)
        return PVL_OK;
    ProfScope prof("mlp_forward", (double)batch);
    if (int st = forward(h, X_dev, batch))
        return st;
    PVL_HIP(hipMemcpyAsync(Y_dev, h->d_z[h->n_layers - 1], sizeof(float) * batch, hipMemcpyDeviceToDevice,
                           stream()));
    return PVL_OK;
}

// Gradients of the mean squared error against T (no update): grads_host (n_params FP32).
int pvl_mlp_gradient(pvl_mlp *h, const float *X_dev, const float *T_dev, int batch, float *grads_host,
                     float *loss_host)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (batch <= 0)
    {
        set_error("mlp_gradient: batch %d", batch);
        return PVL_ERR_ARG;
    }
    if (int st = forward(h, X_dev, batch))
        return st;
    if (int st = backward(h, T_dev, batch))
        return st;
    if (grads_host)
        PVL_HIP(hipMemcpyAsync(grads_host, h->d_g, sizeof(float) * h->n_params, hipMemcpyDeviceToHost, stream()));
    if (loss_host)
        PVL_HIP(hipMemcpyAsync(loss_host, h->d_loss, sizeof(float), hipMemcpyDeviceToHost, stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    return PVL_OK;
}

// One Adam step on a batch, with L2 weight decay (wd * p added to the gradient); loss_host
// (optional) receives the batch loss before the update (synchronises only if requested).
int pvl_mlp_train_step(pvl_mlp *h, const float *X_dev, const float *T_dev, int batch, float lr, float beta1,
                       float beta2, float eps, float wd, float *loss_host)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (batch <= 0)
    {
        set_error("mlp_train_step: batch %d", batch);
        return PVL_ERR_ARG;
    }
    ProfScope prof("mlp_train_step", (double)batch);
    if (int st = forward(h, X_dev, batch))
        return st;
    if (int st = backward(h, T_dev, batch))
        return st;
    h->step += 1;
    const float c1 = 1.0f - std::pow(beta1, (float)h->step), c2 = 1.0f - std::pow(beta2, (float)h->step);
    adam_kernel<<<256, 256, 0, stream()>>>(h->d_p, h->d_g, h->d_m, h->d_v, h->d_pb, (long)h->n_params, lr, beta1,
                                           beta2, eps, c1, c2, wd);
    PVL_HIP(hipGetLastError());
    if (loss_host)
    {
        PVL_HIP(hipMemcpyAsync(loss_host, h->d_loss, sizeof(float), hipMemcpyDeviceToHost, stream()));
        PVL_HIP(hipStreamSynchronize(stream()));
    }
    return PVL_OK;
}

} // extern "C"
