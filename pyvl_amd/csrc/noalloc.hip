// pvl_point_velocity_noalloc: the point-velocity sum of pvl_point_velocity for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  The same kernels with the same chunking, so results are
// bit-identical to pvl_point_velocity; only the working buffers differ: line geometry, line
// circulations and per-chunk partial sums live in one fixed per-device array instead of the
// library's grown-on-demand scratch.
#include "pvl_internal.hpp"

namespace
{

constexpr size_t kScratchBytes = PVL_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

// Device address of `scratch` on each device, from hipGetSymbolAddress at set-up.
void *g_scratch_dev[kMaxDevices] = {};

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

} // namespace

// Working memory of the no-allocation entry point, one copy per device, held until the
// process ends.  In host code this name is only a host-side stand-in: kernels receive the
// per-device address kept in g_scratch_dev, never this symbol.  Nothing in it needs to start
// at zero: every call writes the geometry, circulations and partial sums it then reads.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace pvl
{

int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
    {
        set_error("device ordinal %d beyond the %d the no-allocation entry point tracks", device, kMaxDevices);
        return PVL_ERR_ARG;
    }
    if (g_scratch_dev[device])
        return PVL_OK;
    void *p = nullptr;
    // ::scratch, the device array: inside namespace pvl the bare name is pvl::scratch().
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (e != hipSuccess)
    {
        (void)hipGetLastError(); // do not leave the failure to be reported by a later launch
        return check_hip(e, "hipGetSymbolAddress(scratch)");
    }
    if (!p)
    {
        set_error("hipGetSymbolAddress(scratch) returned a null address on device %d", device);
        return PVL_ERR_HIP;
    }
    g_scratch_dev[device] = p;
    return PVL_OK;
}

} // namespace pvl

using namespace pvl;

extern "C" int pvl_point_velocity_noalloc(const pvl_mesh *m, double tol, const double *pos, const double *circ,
                                          unsigned n_pts, const double *pts, double *out, int accumulate,
                                          int precision)
{
    PVL_API_LOCK;
    if (!m)
    {
        set_error("null mesh");
        return PVL_ERR_ARG;
    }
    if (precision != PVL_FP64 && precision != PVL_FP32 && precision != PVL_FP64_REF)
    {
        set_error("unknown precision %d", precision);
        return PVL_ERR_ARG;
    }
    const int dev = initialised_device();
    if (dev < 0)
    {
        set_error("pvl_point_velocity_noalloc: call pvl_init first (set-up creates the library stream)");
        return PVL_ERR_ARG;
    }
    if (int st = ensure_init()) // makes the library's device current; allocates nothing once initialised
        return st;
    if (int st = noalloc_setup(dev)) // done at initialisation; repeated here only if that failed
        return st;
    if (n_pts == 0)
        return PVL_OK;
    const unsigned n_active = m->n_lines;
    if (n_active == 0)
        return accumulate ? PVL_OK : device_fill32(out, 0u, sizeof(double) * 3 * (size_t)n_pts);

    const unsigned n_chunks = point_velocity_chunks(n_pts, n_active, "point_velocity");
    const size_t geom_bytes = align256(line_geom_bytes(n_active));
    const size_t gamma_bytes = align256(sizeof(double) * (size_t)n_active);
    const size_t partial_bytes = sizeof(double) * 3 * (size_t)n_pts * n_chunks;
    const size_t need = geom_bytes + gamma_bytes + partial_bytes;
    if (need > kScratchBytes)
    {
        set_error("pvl_point_velocity_noalloc: %u points x %u lines need %zu bytes of scratch, more than its fixed "
                  "%zu (use pvl_point_velocity)",
                  n_pts, n_active, need, kScratchBytes);
        return PVL_ERR_ARG;
    }
    unsigned char *base = (unsigned char *)g_scratch_dev[dev];
    void *geom = base;
    double *gamma = (double *)(base + geom_bytes);
    double *partial = (double *)(base + geom_bytes + gamma_bytes);
    if (int st = point_velocity_partials(m, tol, pos, circ, n_pts, pts, precision, n_active, "point_velocity",
                                         stream(), geom, gamma, partial, 0))
        return st;
    return point_velocity_reduce(n_pts, n_chunks, partial, out, accumulate, stream());
}
