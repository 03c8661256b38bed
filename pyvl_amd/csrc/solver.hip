// Small kernels of the run_solver step (python/pyvl/solver.py:175-239) that keep the step
// resident on the device.
#include "pvl_internal.hpp"

using namespace pvl;

namespace
{

// b_i = n_i . -(V_inf(cp_i) - v_cp,i)  (python/pyvl/solver.py:199-206; dot evaluated left to
// right).  cpts is read only for a non-uniform field.
__global__ void rhs_kernel(unsigned n, const double *__restrict__ normals, const double *__restrict__ cpvel,
                           const double *__restrict__ cpts, FlowField flow, double *__restrict__ out)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double *nn = normals + 3 * i;
    double vx, vy, vz;
    if (flow.kind == 0)
        flow_velocity(flow, 0.0, 0.0, 0.0, vx, vy, vz);
    else
        flow_velocity(flow, cpts[3 * i + 0], cpts[3 * i + 1], cpts[3 * i + 2], vx, vy, vz);
    const double e0 = -(vx - (cpvel ? cpvel[3 * i + 0] : 0.0));
    const double e1 = -(vy - (cpvel ? cpvel[3 * i + 1] : 0.0));
    const double e2 = -(vz - (cpvel ? cpvel[3 * i + 2] : 0.0));
    out[i] = nn[0] * e0 + nn[1] * e1 + nn[2] * e2;
}

// x[r0:r1] -= mean(x[r0:r1]) in one workgroup (python/pyvl/solver.py:229-233)
__global__ void __launch_bounds__(1024) remove_mean_kernel(double *__restrict__ x, int r0, int r1)
{
    main_stream_priority();
    __shared__ double part[1024];
    double s = 0.0;
    for (int i = r0 + threadIdx.x; i < r1; i += 1024)
        s += x[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    const double mean = part[0] / (double)(r1 - r0);
    for (int i = r0 + threadIdx.x; i < r1; i += 1024)
        x[i] = x[i] - mean;
}

} // namespace

int flow_from_descriptor(const double *desc, FlowField *out)
{
    if (!desc)
    {
        set_error("null free-stream descriptor");
        return PVL_ERR_ARG;
    }
    if (desc[0] != 0.0 && desc[0] != 1.0)
    {
        set_error("unknown free-stream kind %g (0: uniform, 1: rotating)", desc[0]);
        return PVL_ERR_ARG;
    }
    out->kind = (int)desc[0];
    for (int k = 0; k < 3; ++k)
        out->a[k] = desc[1 + k], out->b[k] = desc[4 + k];
    return PVL_OK;
}

extern "C" {

int pvl_rhs_flow(unsigned n, const double *normals_dev, const double *cpvel_dev, const double *cpts_dev,
                 const double *flow, double *out_dev)
{
    PVL_API_LOCK;
    FlowField f;
    if (int st = flow_from_descriptor(flow, &f))
        return st;
    if (f.kind != 0 && !cpts_dev)
    {
        set_error("a non-uniform free stream needs the control points");
        return PVL_ERR_ARG;
    }
    if (int st = ensure_init())
        return st;
    if (n == 0)
        return PVL_OK;
    rhs_kernel<<<(n + 255) / 256, 256, 0, stream()>>>(n, normals_dev, cpvel_dev, cpts_dev, f, out_dev);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

int pvl_rhs_normal_velocity(unsigned n, const double *normals_dev, const double *cpvel_dev, const double *vinf,
                            double *out_dev)
{
    const double desc[7] = {0.0, vinf[0], vinf[1], vinf[2], 0.0, 0.0, 0.0};
    return pvl_rhs_flow(n, normals_dev, cpvel_dev, nullptr, desc, out_dev);
}

int pvl_remove_mean(double *x_dev, int r0, int r1)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (r0 < 0 || r1 <= r0)
    {
        set_error("remove_mean: bad range [%d, %d)", r0, r1);
        return PVL_ERR_ARG;
    }
    remove_mean_kernel<<<1, 1024, 0, stream()>>>(x_dev, r0, r1);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

} // extern "C"
