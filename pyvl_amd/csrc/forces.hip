// Post-processing on the GPU: Kutta-Joukowski line forces (Mesh.line_forces).
//
// One lane per primal line.  The dual mesh supplies, per line, the surface on its + side
// (dual line start) and on its - side (dual line end); INVALID_ID ends contribute nothing.
// Compiled with -ffp-contract=off: every expression rounds as written, in the operation order
// of the reference's real3_add / real3_mul1 / real3_cross (src/core/common.h:98-136), so the
// forces equal the reference's bit for bit on the same inputs.
#include "pvl_internal.hpp"

using namespace pvl;

namespace
{

// dual_lines: the dual mesh's lines (surface on the + side, surface on the - side; INVALID_ID
// where a line has one surface); an index beyond n_surfaces is treated as absent.
__global__ void __launch_bounds__(256)
    line_forces_kernel(unsigned n_lines, unsigned n_surfaces, const uint32_t *__restrict__ primal_lines,
                       const uint32_t *__restrict__ dual_lines, const double *__restrict__ circ,
                       const double *__restrict__ pos, const double *__restrict__ vel, double *__restrict__ out)
{
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_lines)
        return;
    const uint32_t p1 = primal_lines[2 * i] & kInvalidId, p2 = primal_lines[2 * i + 1] & kInvalidId;
    const uint32_t s_pos = dual_lines[2 * i] & kInvalidId, s_neg = dual_lines[2 * i + 1] & kInvalidId;
    const double drx = pos[3 * p2 + 0] - pos[3 * p1 + 0];
    const double dry = pos[3 * p2 + 1] - pos[3 * p1 + 1];
    const double drz = pos[3 * p2 + 2] - pos[3 * p1 + 2];
    double c = 0.0;
    if (s_pos < n_surfaces)
        c = c + circ[s_pos];
    if (s_neg < n_surfaces)
        c = c - circ[s_neg];
    const double k = 0.5 * c;
    const double ax = (vel[3 * p1 + 0] + vel[3 * p2 + 0]) * k;
    const double ay = (vel[3 * p1 + 1] + vel[3 * p2 + 1]) * k;
    const double az = (vel[3 * p1 + 2] + vel[3 * p2 + 2]) * k;
    out[3 * i + 0] = dry * az - drz * ay;
    out[3 * i + 1] = drz * ax - drx * az;
    out[3 * i + 2] = drx * ay - dry * ax;
}

} // namespace

extern "C" {

int pvl_line_forces(const pvl_mesh *primal, const uint32_t *dual_lines_dev, unsigned dual_n_points,
                    unsigned dual_n_lines, unsigned dual_n_surfaces, const double *circulation_dev,
                    const double *positions_dev, const double *velocity_dev, double *out_dev)
{
    PVL_API_LOCK;
    if (!primal)
    {
        set_error("null mesh");
        return PVL_ERR_ARG;
    }
    if (primal->n_points != dual_n_surfaces || primal->n_lines != dual_n_lines || primal->n_surfaces != dual_n_points)
    {
        set_error("Given meshes can not be dual to each other, since the number of points,lines, and surfaces don't "
                  "match as primal (%u, %u, %u) and dual (%u, %u, %u).",
                  primal->n_points, primal->n_lines, primal->n_surfaces, dual_n_points, dual_n_lines,
                  dual_n_surfaces);
        return PVL_ERR_ARG;
    }
    if (int st = ensure_init())
        return st;
    const unsigned n = primal->n_lines;
    if (n == 0)
        return PVL_OK;
    ProfScope prof("line_forces", (double)n);
    line_forces_kernel<<<(n + 255) / 256, 256, 0, stream()>>>(n, primal->n_surfaces, primal->d_lines, dual_lines_dev,
                                                             circulation_dev, positions_dev, velocity_dev, out_dev);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

} // extern "C"
