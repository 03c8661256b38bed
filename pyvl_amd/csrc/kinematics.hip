// Rigid-frame kinematics of the time step on the GPU (python/pyvl/solver.py:180-193).
//
// The reference walks each body's reference-frame chain on the host, child to root:
//   positions  = rf.to_parent_with_offset(positions)      R^T (p - d)
//   rf.add_velocity(positions, velocities)                v += frame term at the new p
//   velocities = rf.to_parent_without_offset(velocities)  R^T v
// Here the frames are evaluated at the step's time once on the host (their at_time, the
// rotation matrices and velocity terms: a few dozen numbers per frame) and passed by value;
// one kernel applies the whole chain to every point of the body, which stays in HBM.
// Velocity terms: none, constant (translating frames), p x w (rotor frames), W p + c
// (oscillating frames).  Matrix-vector products are the reference's real3x3_vecmul,
// (m0 x + m1 y) + m2 z, and with -ffp-contract=off every expression rounds as written.
#include "pvl_internal.hpp"

using namespace pvl;

namespace
{

constexpr int kMaxFrames = 8;
constexpr int kTerms = 24; // rinv[9], offset[3], w[9], c[3]

struct FrameChain
{
    double t[kMaxFrames][kTerms];
    int kind[kMaxFrames];
    unsigned n;
};

__device__ __forceinline__ void vecmul(const double *m, double x, double y, double z, double &ox, double &oy,
                                       double &oz)
{
    ox = (m[0] * x + m[1] * y) + m[2] * z;
    oy = (m[3] * x + m[4] * y) + m[5] * z;
    oz = (m[6] * x + m[7] * y) + m[8] * z;
}

__global__ void __launch_bounds__(256) frame_kinematics_kernel(unsigned n_points, const double *__restrict__ p0,
                                                              const FrameChain chain, double *__restrict__ pos,
                                                              double *__restrict__ vel)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_points)
        return;
    double px = p0[3 * i + 0], py = p0[3 * i + 1], pz = p0[3 * i + 2];
    double vx = 0.0, vy = 0.0, vz = 0.0;
    for (unsigned f = 0; f < chain.n; ++f)
    {
        const double *t = chain.t[f];
        const double *rinv = t, *off = t + 9, *w = t + 12, *c = t + 21;
        vecmul(rinv, px - off[0], py - off[1], pz - off[2], px, py, pz);
        switch (chain.kind[f])
        {
        case PVL_FRAME_TRANSLATING:
            vx = vx + c[0];
            vy = vy + c[1];
            vz = vz + c[2];
            break;
        case PVL_FRAME_ROTOR: // v += p x omega (numpy.cross order)
            vx = vx + (py * c[2] - pz * c[1]);
            vy = vy + (pz * c[0] - px * c[2]);
            vz = vz + (px * c[1] - py * c[0]);
            break;
        case PVL_FRAME_AFFINE:
        {
            double ax, ay, az;
            vecmul(w, px, py, pz, ax, ay, az);
            vx = vx + (ax + c[0]);
            vy = vy + (ay + c[1]);
            vz = vz + (az + c[2]);
            break;
        }
        default:
            break;
        }
        vecmul(rinv, vx, vy, vz, vx, vy, vz);
    }
    pos[3 * i + 0] = px;
    pos[3 * i + 1] = py;
    pos[3 * i + 2] = pz;
    if (vel)
    {
        vel[3 * i + 0] = vx;
        vel[3 * i + 1] = vy;
        vel[3 * i + 2] = vz;
    }
}

} // namespace

extern "C" int pvl_frame_kinematics(unsigned n_points, const double *p0_dev, unsigned n_frames, const double *terms,
                                    const int *kinds, double *pos_dev, double *vel_dev)
{
    PVL_API_LOCK;
    if (n_frames > (unsigned)kMaxFrames)
    {
        set_error("frame chain of %u frames is longer than the %d supported", n_frames, kMaxFrames);
        return PVL_ERR_ARG;
    }
    if (n_frames && (!terms || !kinds))
    {
        set_error("null frame terms");
        return PVL_ERR_ARG;
    }
    FrameChain chain{};
    chain.n = n_frames;
    for (unsigned f = 0; f < n_frames; ++f)
    {
        if (kinds[f] < PVL_FRAME_NONE || kinds[f] > PVL_FRAME_AFFINE)
        {
            set_error("unknown frame kind %d", kinds[f]);
            return PVL_ERR_ARG;
        }
        chain.kind[f] = kinds[f];
        for (int k = 0; k < kTerms; ++k)
            chain.t[f][k] = terms[(size_t)f * kTerms + k];
    }
    if (int st = ensure_init())
        return st;
    if (n_points == 0)
        return PVL_OK;
    frame_kinematics_kernel<<<(n_points + 255) / 256, 256, 0, stream()>>>(n_points, p0_dev, chain, pos_dev, vel_dev);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}
