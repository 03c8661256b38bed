// Line-vortex induction on gfx950: dense matrix assembly, induction tensors and
// point-velocity sums.  All kernels are compiled with -ffp-contract=off, so every FP64
// expression below rounds exactly as written (no FMA contraction), matching the
// reference's x86-64 baseline build.
#include "pvl_internal.hpp"
#include "induction_core.hpp"

#include <cstring>
#include <type_traits>

// Calls f(std::integral_constant<int, P>{}) for the precision form P selected at run time.
template <class F> static void dispatch_precision(int precision, F &&f)
{
    if (precision == PVL_FP32)
        f(std::integral_constant<int, PVL_FP32>{});
    else if (precision == PVL_FP64_REF_FORM)
        f(std::integral_constant<int, PVL_FP64_REF_FORM>{});
    else
        f(std::integral_constant<int, PVL_FP64>{});
}

using namespace pvl;

namespace
{

constexpr int kWave = 64;
constexpr int kBlock = 256; // 4 waves
constexpr int kCptTile = 64; // control points per workgroup (one per lane)

// ---------------------------------------------------------------------------------------
// Per-line prologue (src/core/flow_solver.c:95-103): end points, |d| and d/|d| with three
// true divisions.
// A line shorter than tol is cut off for every point (K = 0 in every form); its direction
// and length are stored as 0 (instead of the 0/0 of a zero-length line), so that its angle
// and the fused accumulation (0 * direction) are 0.  With tol <= 0 nothing is cut by length
// and the reference's values, NaN included, are kept.
__global__ void __launch_bounds__(256) line_geom_kernel(unsigned n_lines, const uint32_t *__restrict__ lines,
                                                       const double *__restrict__ pos, double tol,
                                                       LineGeom *__restrict__ out)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_lines)
        return;
    const uint32_t p1 = lines[2 * i] & kInvalidId, p2 = lines[2 * i + 1] & kInvalidId;
    LineGeom g;
    g.r1x = pos[3 * p1 + 0], g.r1y = pos[3 * p1 + 1], g.r1z = pos[3 * p1 + 2];
    g.r2x = pos[3 * p2 + 0], g.r2y = pos[3 * p2 + 1], g.r2z = pos[3 * p2 + 2];
    const double dx = g.r2x - g.r1x, dy = g.r2y - g.r1y, dz = g.r2z - g.r1z;
    g.len = sqrt(dx * dx + dy * dy + dz * dz);
    g.ex = dx / g.len;
    g.ey = dy / g.len;
    g.ez = dz / g.len;
    if (g.len < tol)
        g.ex = g.ey = g.ez = g.len = 0.0;
    out[i] = g;
}

__device__ __forceinline__ unsigned wave_id() { return __builtin_amdgcn_readfirstlane(threadIdx.x / kWave); }
__device__ __forceinline__ unsigned lane_id() { return threadIdx.x % kWave; }

// ---------------------------------------------------------------------------------------
// Surface-reduced induction tensor: one lane per control point, the wave walks a run of
// surfaces (wave-uniform, so surface/line metadata and line geometry come through scalar
// loads).  Each surface accumulates +-K in its CSR order (src/core/flow_solver.c:141-172).
// K of the lines of the previous surface is kept in registers: consecutive surfaces of
// structured meshes share a line.  Results are staged in LDS and written row-contiguous.
//   VEC:    out[n_cpts][n_surf][3]   (induction_matrix)
//   NORMAL: out[n_cpts][n_surf], dotted with the control point's normal (:202); the library
//           assembles induction_matrix3 with assembly_kernel instead.
template <bool NORMAL, int PREC, int SURF_PER_WAVE>
__global__ void __launch_bounds__(kBlock)
    surface_induction_kernel(unsigned n_cpts, unsigned n_surf, const uint32_t *__restrict__ offsets,
                             const uint32_t *__restrict__ entries, const LineGeom *__restrict__ geom,
                             const double *__restrict__ cpts, const double *__restrict__ normals,
                             double *__restrict__ out, double tol)
{
    main_stream_priority();
    constexpr int kSurfTile = 4 * SURF_PER_WAVE;
    constexpr int kComp = NORMAL ? 1 : 3;
    constexpr int kRow = kSurfTile * kComp + 1; // padded LDS row (doubles)
    __shared__ double tile[kCptTile * kRow];

    const unsigned w = wave_id(), lane = lane_id();
    const unsigned cp0 = blockIdx.y * kCptTile;
    const unsigned s_tile0 = blockIdx.x * kSurfTile;
    const unsigned cp = cp0 + lane;
    const bool live = cp < n_cpts;
    const unsigned cpc = live ? cp : n_cpts - 1;
    const double px = cpts[3 * cpc + 0], py = cpts[3 * cpc + 1], pz = cpts[3 * cpc + 2];
    double nx = 0, ny = 0, nz = 0;
    if (NORMAL)
        nx = normals[3 * cpc + 0], ny = normals[3 * cpc + 1], nz = normals[3 * cpc + 2];
    const double tol2 = tol * tol;

    // register cache of the previous surface's first four lines
    uint32_t oid0 = 0xFFFFFFFFu, oid1 = 0xFFFFFFFFu, oid2 = 0xFFFFFFFFu, oid3 = 0xFFFFFFFFu;
    double o0x = 0, o0y = 0, o0z = 0, o1x = 0, o1y = 0, o1z = 0, o2x = 0, o2y = 0, o2z = 0, o3x = 0, o3y = 0,
           o3z = 0;

    for (int j = 0; j < SURF_PER_WAVE; ++j)
    {
        const unsigned sl = w * SURF_PER_WAVE + j; // local surface
        const unsigned s = s_tile0 + sl;
        double rx = 0.0, ry = 0.0, rz = 0.0;
        uint32_t nid0 = 0xFFFFFFFFu, nid1 = 0xFFFFFFFFu, nid2 = 0xFFFFFFFFu, nid3 = 0xFFFFFFFFu;
        double c0x = 0, c0y = 0, c0z = 0, c1x = 0, c1y = 0, c1z = 0, c2x = 0, c2y = 0, c2z = 0, c3x = 0, c3y = 0,
               c3z = 0;
        if (s < n_surf)
        {
            const uint32_t k0 = offsets[s], k1 = offsets[s + 1];
            for (uint32_t k = k0; k < k1; ++k)
            {
                const uint32_t e = entries[k];
                const uint32_t l = e & kInvalidId;
                double kx, ky, kz;
                if (l == oid0)
                    kx = o0x, ky = o0y, kz = o0z;
                else if (l == oid1)
                    kx = o1x, ky = o1y, kz = o1z;
                else if (l == oid2)
                    kx = o2x, ky = o2y, kz = o2z;
                else if (l == oid3)
                    kx = o3x, ky = o3y, kz = o3z;
                else
                    line_k<PREC>(geom[l], px, py, pz, tol, tol2, kx, ky, kz);
                const uint32_t slot = k - k0;
                if (slot == 0)
                    nid0 = l, c0x = kx, c0y = ky, c0z = kz;
                else if (slot == 1)
                    nid1 = l, c1x = kx, c1y = ky, c1z = kz;
                else if (slot == 2)
                    nid2 = l, c2x = kx, c2y = ky, c2z = kz;
                else if (slot == 3)
                    nid3 = l, c3x = kx, c3y = ky, c3z = kz;
                if (e & kReversed)
                {
                    rx = rx - kx;
                    ry = ry - ky;
                    rz = rz - kz;
                }
                else
                {
                    rx = rx + kx;
                    ry = ry + ky;
                    rz = rz + kz;
                }
            }
        }
        oid0 = nid0, oid1 = nid1, oid2 = nid2, oid3 = nid3;
        o0x = c0x, o0y = c0y, o0z = c0z, o1x = c1x, o1y = c1y, o1z = c1z;
        o2x = c2x, o2y = c2y, o2z = c2z, o3x = c3x, o3y = c3y, o3z = c3z;
        if (NORMAL)
        {
            tile[lane * kRow + sl] = rx * nx + ry * ny + rz * nz;
        }
        else
        {
            tile[lane * kRow + 3 * sl + 0] = rx;
            tile[lane * kRow + 3 * sl + 1] = ry;
            tile[lane * kRow + 3 * sl + 2] = rz;
        }
    }
    __syncthreads();
    // Row-contiguous store: each wave writes 16 control-point rows.
    const unsigned s_count = (n_surf - s_tile0) < kSurfTile ? (n_surf - s_tile0) : kSurfTile;
    const unsigned row_len = s_count * kComp;
    for (unsigned r = w; r < kCptTile; r += 4)
    {
        const unsigned cpr = cp0 + r;
        if (cpr >= n_cpts)
            break;
        double *dst = out + ((size_t)cpr * n_surf + s_tile0) * kComp;
        for (unsigned c = lane; c < row_len; c += kWave)
            dst[c] = tile[r * kRow + c];
    }
}

// Normal-reduced assembly (induction_matrix3), out[n_cpts][n_surf]: one lane per control
// point, one wave per run of up to kAsmGroups groups of kAsmSurfaces consecutive surfaces,
// executing the run's program (build_assembly_program): each line of the run is evaluated
// once, as n . K, and kept in an LDS slot while a later surface of the run needs it; a
// surface's value is the sum of its lines' +-n.K in CSR order (src/core/flow_solver.c:188-202
// sums the vectors first and then dots with the normal: the same terms, another association).
// Program entries are wave-uniform (scalar loads and branches); per lane the inner step is one
// K and one LDS access at most.  Each group's results are staged in LDS and stored as rows of
// kAsmSurfaces consecutive doubles.
template <int PREC>
__global__ void __launch_bounds__(kBlock)
    assembly_kernel(unsigned n_cpts, unsigned n_surf, unsigned n_runs, const uint32_t *__restrict__ run_off,
                    const uint32_t *__restrict__ run_grp, const uint32_t *__restrict__ prog,
                    const LineGeom *__restrict__ geom, const double *__restrict__ cpts,
                    const double *__restrict__ normals, double *__restrict__ out, double tol)
{
    main_stream_priority();
    __shared__ double slots[4][kAsmSlots][kWave];
    __shared__ double res[4][kAsmSurfaces][kWave + 1];
    const unsigned w = wave_id(), lane = lane_id();
    const unsigned cp0 = blockIdx.y * kCptTile;
    const unsigned cp = cp0 + lane;
    const unsigned cpc = cp < n_cpts ? cp : n_cpts - 1;
    const double px = cpts[3 * cpc + 0], py = cpts[3 * cpc + 1], pz = cpts[3 * cpc + 2];
    const double nx = normals[3 * cpc + 0], ny = normals[3 * cpc + 1], nz = normals[3 * cpc + 2];
    const double tol2 = tol * tol;
    const unsigned run = blockIdx.x * 4 + w;
    if (run >= n_runs)
        return;
    const uint32_t k0 = run_off[run], k1 = run_off[run + 1];
    unsigned gi = 0;
    uint32_t grp = run_grp[run * kAsmGroups];
    unsigned ns = min(kAsmSurfaces, n_surf - grp * kAsmSurfaces); // surfaces of the current group
    unsigned sl = 0;
    double acc = 0.0;
    // pairs of entries: both entries' scalar loads (program words and line geometry) are
    // issued before either is evaluated, so one wait covers two evaluations (scalar loads
    // return out of order: a wave can only wait for all of them)
    auto step = [&](uint32_t e, uint32_t op, const LineGeom &g) {
        const uint32_t kind = op & 3u, slot = (op >> 2) & 15u;
        double v;
        if (kind == kAsmLoad)
            v = slots[w][slot][lane];
        else if (kind == kAsmNone)
            v = 0.0;
        else
        {
            v = line_nk<PREC>(g, px, py, pz, nx, ny, nz, tol, tol2);
            if (kind == kAsmEvalStore)
                slots[w][slot][lane] = v;
        }
        acc = (e & kReversed) ? acc - v : acc + v;
        if (op & kAsmEnd)
        {
            res[w][sl][lane] = acc;
            acc = 0.0;
            ++sl;
        }
    };
    // groups have even entry counts (padded), so a group ends on a pair boundary
    for (uint32_t k = k0; k < k1; k += 2)
    {
        // a kAsmNone entry names line 0 (the geometry buffer always has one element)
        const uint32_t ea = prog[2 * k], opa = prog[2 * k + 1], eb = prog[2 * k + 2], opb = prog[2 * k + 3];
        const LineGeom ga = geom[ea & kInvalidId], gb = geom[eb & kInvalidId];
        // both geometries in SGPRs here (the compiler would otherwise sink the second load
        // into the branch that uses it, after the first evaluation)
        asm volatile("" ::"s"(ga.r1x), "s"(ga.r2x), "s"(ga.ex), "s"(ga.len), "s"(gb.r1x), "s"(gb.r2x), "s"(gb.ex),
                     "s"(gb.len));
        step(ea, opa, ga);
        step(eb, opb, gb);
        if (sl == ns)
        {
            // the group's rows: lane = 8 rows x 8 surfaces per store instruction
            const unsigned s0 = grp * kAsmSurfaces, c = lane % kAsmSurfaces;
            for (unsigned r = lane / kAsmSurfaces; r < kCptTile; r += kWave / kAsmSurfaces)
            {
                const unsigned p = cp0 + r;
                if (p < n_cpts && c < ns)
                    out[(size_t)p * n_surf + s0 + c] = res[w][c][r];
            }
            sl = 0;
            if (++gi < kAsmGroups)
            {
                grp = run_grp[run * kAsmGroups + gi];
                if (grp != kAsmNoGroup)
                    ns = min(kAsmSurfaces, n_surf - grp * kAsmSurfaces);
            }
        }
    }
}

// Raw line tensor out[n_cpts][n_lines][3] (src/core/flow_solver.c:86-139 with its
// cpt-major layout :110,136).
template <int PREC, int LINES_PER_WAVE>
__global__ void __launch_bounds__(kBlock)
    line_induction_kernel(unsigned n_cpts, unsigned n_lines, const LineGeom *__restrict__ geom,
                          const double *__restrict__ cpts, double *__restrict__ out, double tol)
{
    main_stream_priority();
    constexpr int kLineTile = 4 * LINES_PER_WAVE;
    constexpr int kRow = kLineTile * 3 + 1;
    __shared__ double tile[kCptTile * kRow];
    const unsigned w = wave_id(), lane = lane_id();
    const unsigned cp0 = blockIdx.y * kCptTile;
    const unsigned l_tile0 = blockIdx.x * kLineTile;
    const unsigned cp = cp0 + lane;
    const unsigned cpc = cp < n_cpts ? cp : n_cpts - 1;
    const double px = cpts[3 * cpc + 0], py = cpts[3 * cpc + 1], pz = cpts[3 * cpc + 2];
    const double tol2 = tol * tol;
    for (int j = 0; j < LINES_PER_WAVE; ++j)
    {
        const unsigned ll = w * LINES_PER_WAVE + j;
        const unsigned l = l_tile0 + ll;
        double kx = 0, ky = 0, kz = 0;
        if (l < n_lines)
            line_k<PREC>(geom[l], px, py, pz, tol, tol2, kx, ky, kz);
        tile[lane * kRow + 3 * ll + 0] = kx;
        tile[lane * kRow + 3 * ll + 1] = ky;
        tile[lane * kRow + 3 * ll + 2] = kz;
    }
    __syncthreads();
    const unsigned l_count = (n_lines - l_tile0) < kLineTile ? (n_lines - l_tile0) : kLineTile;
    const unsigned row_len = 3 * l_count;
    for (unsigned r = w; r < kCptTile; r += 4)
    {
        const unsigned cpr = cp0 + r;
        if (cpr >= n_cpts)
            break;
        double *dst = out + ((size_t)cpr * n_lines + l_tile0) * 3;
        for (unsigned c = lane; c < row_len; c += kWave)
            dst[c] = tile[r * kRow + c];
    }
}

// ---------------------------------------------------------------------------------------
// Point-velocity sums.  Line circulation gamma_l = sum over the surfaces holding line l
// of +-Gamma_s, in increasing surface order (deterministic).
__global__ void __launch_bounds__(256) line_circulation_kernel(unsigned n_lines, const uint32_t *__restrict__ loff,
                                                               const uint32_t *__restrict__ lsurf,
                                                               const double *__restrict__ circ,
                                                               double *__restrict__ gamma)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_lines)
        return;
    double g = 0.0;
    for (uint32_t k = loff[i]; k < loff[i + 1]; ++k)
    {
        const uint32_t e = lsurf[k];
        const double c = circ[e & kInvalidId];
        g = (e & kReversed) ? g - c : g + c;
    }
    gamma[i] = g;
}

// PPL points per lane (64 PPL per wave, point j * 64 + lane of the block); the workgroup's
// four waves split a chunk of lines, partial sums are combined in LDS in wave order and
// written per chunk; a second pass adds the chunks in chunk order.  No atomics: the result
// does not depend on scheduling.  Work items are (64 PPL-point block, line chunk) pairs,
// item = chunk * pblocks + block; workgroup b takes items b, b + gridDim.x, .. (a grid
// smaller than the item count leaves CU slots to other work on the device; each item's
// arithmetic is the same either way).  With PPL = 2 each lane runs two independent
// evaluation chains per line (instruction-level parallelism) and every line's geometry,
// loaded once into scalar registers, serves twice the points.
// SITE only gives the free-wake self-induction launch its own symbol for profilers; it
// uses PPL = 2, the other sites PPL = 1.  The self-induction kernel needs 60 VGPRs: five
// workgroups per CU on the side stream (one wave each per SIMD) leave each SIMD 192 VGPRs
// for the library stream's kernels.
template <int SITE> constexpr int site_ppl() { return SITE == 1 ? 2 : 1; }

template <int PREC, int SITE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(SITE == 1 ? 4 : 6)))
    point_velocity_kernel(unsigned n_pts, const double *__restrict__ pts, unsigned n_lines,
                          const LineGeom *__restrict__ geom, const double *__restrict__ gamma,
                          unsigned lines_per_chunk, double *__restrict__ partial, double tol, unsigned pblocks,
                          unsigned n_items)
{
    if constexpr (SITE == 0)
        main_stream_priority();
    constexpr int PPL = site_ppl<SITE>();
    __shared__ double red[4][3][kWave * PPL];
    const unsigned w = wave_id(), lane = lane_id();
    const double tol2 = tol * tol;
    const unsigned per_wave = (lines_per_chunk + 3) / 4;
    for (unsigned item = blockIdx.x; item < n_items; item += gridDim.x)
    {
        const unsigned chunk = item / pblocks;
        const unsigned p0 = (item - chunk * pblocks) * (kWave * PPL) + lane;
        double px[PPL], py[PPL], pz[PPL], vx[PPL], vy[PPL], vz[PPL];
#pragma unroll
        for (int j = 0; j < PPL; ++j)
        {
            const unsigned p = p0 + j * kWave;
            const unsigned pc = p < n_pts ? p : n_pts - 1;
            px[j] = pts[3 * pc + 0], py[j] = pts[3 * pc + 1], pz[j] = pts[3 * pc + 2];
            vx[j] = 0.0, vy[j] = 0.0, vz[j] = 0.0;
        }
        const unsigned l0 = chunk * lines_per_chunk;
        const unsigned l1 = (l0 + lines_per_chunk) < n_lines ? (l0 + lines_per_chunk) : n_lines;
        const unsigned a = l0 + w * per_wave;
        const unsigned b = (a + per_wave) < l1 ? (a + per_wave) : l1;
        for (unsigned l = a; l < b; ++l)
        {
            const LineGeom g = geom[l];
            const double gl = gamma[l];
#pragma unroll
            for (int j = 0; j < PPL; ++j)
                line_k_acc<PREC>(g, px[j], py[j], pz[j], tol, tol2, gl, vx[j], vy[j], vz[j]);
        }
#pragma unroll
        for (int j = 0; j < PPL; ++j)
        {
            red[w][0][j * kWave + lane] = vx[j];
            red[w][1][j * kWave + lane] = vy[j];
            red[w][2][j * kWave + lane] = vz[j];
        }
        __syncthreads();
        if (w == 0)
        {
#pragma unroll
            for (int j = 0; j < PPL; ++j)
            {
                const unsigned p = p0 + j * kWave;
                if (p >= n_pts)
                    continue;
                const unsigned i = j * kWave + lane;
                double sx = red[0][0][i], sy = red[0][1][i], sz = red[0][2][i];
                for (int q = 1; q < 4; ++q)
                {
                    sx = sx + red[q][0][i];
                    sy = sy + red[q][1][i];
                    sz = sz + red[q][2][i];
                }
                double *dst = partial + ((size_t)chunk * n_pts + p) * 3;
                dst[0] = sx;
                dst[1] = sy;
                dst[2] = sz;
            }
        }
        __syncthreads(); // red is rewritten by the next item
    }
}

__global__ void __launch_bounds__(256) chunk_reduce_kernel(unsigned n_pts, unsigned n_chunks,
                                                           const double *__restrict__ partial,
                                                           double *__restrict__ out, int accumulate)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * n_pts)
        return;
    double s = accumulate ? out[i] : 0.0;
    for (unsigned c = 0; c < n_chunks; ++c)
        s = s + partial[(size_t)c * 3 * n_pts + i];
    out[i] = s;
}

// ---------------------------------------------------------------------------------------
// Surface normal and centre (src/core/mesh.c:14-68).
__global__ void __launch_bounds__(256) surface_geometry_kernel(unsigned n_surf, const uint32_t *__restrict__ offsets,
                                                               const uint32_t *__restrict__ entries,
                                                               const uint32_t *__restrict__ lines,
                                                               const double *__restrict__ pos,
                                                               double *__restrict__ normal_out,
                                                               double *__restrict__ center_out)
{
    main_stream_priority();
    const unsigned s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_surf)
        return;
    const uint32_t k0 = offsets[s], k1 = offsets[s + 1];
    if (k1 <= k0)
    {
        // empty surface: no entries to read (the reference divides 0 by 0 here)
        const double nan = __builtin_nan("");
        for (int c = 0; c < 3; ++c)
        {
            if (center_out)
                center_out[3 * s + c] = nan;
            if (normal_out)
                normal_out[3 * s + c] = nan;
        }
        return;
    }
    if (center_out)
    {
        // sum of each line's start point in traversal direction, times 1/n
        double cx = 0.0, cy = 0.0, cz = 0.0;
        for (uint32_t k = k0; k < k1; ++k)
        {
            const uint32_t e = entries[k];
            const uint32_t l = e & kInvalidId;
            const uint32_t p = (e & kReversed) ? (lines[2 * l + 1] & kInvalidId) : (lines[2 * l] & kInvalidId);
            cx = cx + pos[3 * p + 0];
            cy = cy + pos[3 * p + 1];
            cz = cz + pos[3 * p + 2];
        }
        const double div = 1.0 / (double)(k1 - k0);
        center_out[3 * s + 0] = cx * div;
        center_out[3 * s + 1] = cy * div;
        center_out[3 * s + 2] = cz * div;
    }
    if (normal_out)
    {
        // line_direction: p1 - p2 for a forward line, p2 - p1 for a reversed one
        auto dir = [&](uint32_t e, double &x, double &y, double &z) {
            const uint32_t l = e & kInvalidId;
            const uint32_t a = lines[2 * l] & kInvalidId, b = lines[2 * l + 1] & kInvalidId;
            const uint32_t u = (e & kReversed) ? b : a, v = (e & kReversed) ? a : b;
            x = pos[3 * u + 0] - pos[3 * v + 0];
            y = pos[3 * u + 1] - pos[3 * v + 1];
            z = pos[3 * u + 2] - pos[3 * v + 2];
        };
        double d1x, d1y, d1z;
        dir(entries[k0], d1x, d1y, d1z);
        const double d0x = d1x, d0y = d1y, d0z = d1z;
        double ox = 0.0, oy = 0.0, oz = 0.0;
        for (uint32_t k = k0 + 1; k < k1; ++k)
        {
            double d2x, d2y, d2z;
            dir(entries[k], d2x, d2y, d2z);
            // cross(d2, d1)
            ox = ox + (d2y * d1z - d2z * d1y);
            oy = oy + (d2z * d1x - d2x * d1z);
            oz = oz + (d2x * d1y - d2y * d1x);
            d1x = d2x, d1y = d2y, d1z = d2z;
        }
        // closing term cross(d0, d_last), then unit: times 1/|.|
        ox = ox + (d0y * d1z - d0z * d1y);
        oy = oy + (d0z * d1x - d0x * d1z);
        oz = oz + (d0x * d1y - d0y * d1x);
        const double inv = 1.0 / sqrt(ox * ox + oy * oy + oz * oz);
        normal_out[3 * s + 0] = ox * inv;
        normal_out[3 * s + 1] = oy * inv;
        normal_out[3 * s + 2] = oz * inv;
    }
}

// ---------------------------------------------------------------------------------------
int compute_line_geom(const pvl_mesh *m, const double *pos, double tol, LineGeom **out, unsigned n_lines, int slot = 0)
{
    LineGeom *g = (LineGeom *)scratch(slot, sizeof(LineGeom) * (size_t)(n_lines ? n_lines : 1));
    if (!g)
        return PVL_ERR_NOMEM;
    if (n_lines)
    {
        line_geom_kernel<<<(n_lines + 255) / 256, 256, 0, stream()>>>(n_lines, m->d_lines, pos, tol, g);
        PVL_HIP(hipGetLastError());
    }
    *out = g;
    return PVL_OK;
}

int compute_line_geom(const pvl_mesh *m, const double *pos, double tol, LineGeom **out)
{
    return compute_line_geom(m, pos, tol, out, m->n_lines, 0);
}

int check_common(const pvl_mesh *m, double tol, int precision)
{
    if (!m)
    {
        set_error("null mesh");
        return PVL_ERR_ARG;
    }
    if (precision != PVL_FP64 && precision != PVL_FP32 && precision != PVL_FP64_REF_FORM)
    {
        set_error("unknown precision %d", precision);
        return PVL_ERR_ARG;
    }
    (void)tol;
    return ensure_init();
}

template <bool NORMAL>
int launch_surface(const pvl_mesh *m, double tol, const double *pos, unsigned n_cpts, const double *cpts,
                   const double *normals, double *out, int precision)
{
    if (n_cpts == 0 || m->n_surfaces == 0)
        return PVL_OK;
    LineGeom *g;
    if (int st = compute_line_geom(m, pos, tol, &g))
        return st;
    constexpr int spw = 8;
    const dim3 grid(NORMAL ? (m->n_asm_runs + 3) / 4 : (m->n_surfaces + 4 * spw - 1) / (4 * spw),
                    (n_cpts + kCptTile - 1) / kCptTile);
    if (grid.y > 65535u)
    {
        set_error("too many control points (%u) for one launch", n_cpts);
        return PVL_ERR_ARG;
    }
    ProfScope prof(NORMAL ? "induction_matrix3" : "induction_matrix", (double)n_cpts * (double)m->n_lines);
    dispatch_precision(precision, [&](auto P) {
        if constexpr (NORMAL)
            assembly_kernel<decltype(P)::value><<<grid, kBlock, 0, stream()>>>(
                n_cpts, m->n_surfaces, m->n_asm_runs, m->d_asm_off, m->d_asm_grp, m->d_asm, g, cpts, normals, out,
                tol);
        else
            surface_induction_kernel<NORMAL, decltype(P)::value, spw><<<grid, kBlock, 0, stream()>>>(
                n_cpts, m->n_surfaces, m->d_offsets, m->d_entries, g, cpts, normals, out, tol);
    });
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

// ---------------------------------------------------------------------------------------
// Segment chains (the wake's self-induction).  Prologue: each segment's end node, unit
// direction in the chain's direction (three true divisions, the reference's line prologue;
// direction 0 below tol as line_geom_kernel) and circulation, signed for the direction.
__global__ void __launch_bounds__(256) chain_segment_kernel(unsigned n_seg, const uint32_t *__restrict__ seg_ab,
                                                            const uint32_t *__restrict__ seg_line,
                                                            const double *__restrict__ nodes,
                                                            const double *__restrict__ gamma, unsigned n_active,
                                                            double tol, SegRec *__restrict__ out)
{
    const unsigned j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n_seg)
        return;
    const uint32_t a = seg_ab[2 * j], b = seg_ab[2 * j + 1], ln = seg_line[j];
    const double ax = nodes[3 * a + 0], ay = nodes[3 * a + 1], az = nodes[3 * a + 2];
    SegRec r;
    r.nx = nodes[3 * b + 0], r.ny = nodes[3 * b + 1], r.nz = nodes[3 * b + 2];
    const double dx = r.nx - ax, dy = r.ny - ay, dz = r.nz - az;
    r.len = sqrt(dx * dx + dy * dy + dz * dz);
    r.ex = dx / r.len;
    r.ey = dy / r.len;
    r.ez = dz / r.len;
    if (r.len < tol)
        r.ex = r.ey = r.ez = r.len = 0.0;
    const uint32_t l = ln & kInvalidId;
    const double g = l < n_active ? gamma[l] : 0.0;
    r.gamma = (ln & kReversed) ? -g : g;
    out[j] = r;
}

// PPL points per lane; the workgroup's four waves take the pieces of the item's chunk in
// turn; per-wave sums are combined in LDS in wave order and written per chunk (no atomics).
// Along a piece, the end node's dr = p - r becomes the next segment's dr1 (the reference form
// also carries |dr|^2), so each segment record is one 64-byte scalar load and the accurate
// forms spend ~66 VALU instructions per pair (cross_m).  The segment loop is unrolled by two
// so that the carried terms are renamed, not copied.  Capped at 64 VGPRs (a few SGPRs spill to
// VGPR lanes, none to memory): five workgroups per CU on the side stream (five waves per SIMD)
// leave each SIMD 192 VGPRs, room for two 96-VGPR waves of the library stream's kernels.
template <int PREC, int PPL>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8)))
    chain_velocity_kernel(unsigned n_pts, const double *__restrict__ pts, const double *__restrict__ nodes,
                          const ChainPiece *__restrict__ pieces, unsigned n_pieces, unsigned cr,
                          const SegRec *__restrict__ segs, unsigned pieces_per_chunk, double *__restrict__ partial,
                          double tol, unsigned pblocks, unsigned n_items)
{
    __shared__ double red[4][3][kWave * PPL];
    const unsigned w = wave_id(), lane = lane_id();
    const double tol2 = tol * tol;
    for (unsigned item = blockIdx.x; item < n_items; item += gridDim.x)
    {
        const unsigned chunk = item / pblocks;
        const unsigned p0 = (item - chunk * pblocks) * (kWave * PPL) + lane;
        double px[PPL], py[PPL], pz[PPL], vx[PPL], vy[PPL], vz[PPL];
#pragma unroll
        for (int i = 0; i < PPL; ++i)
        {
            const unsigned p = p0 + i * kWave;
            const unsigned pc = p < n_pts ? p : n_pts - 1;
            px[i] = pts[3 * pc + 0], py[i] = pts[3 * pc + 1], pz[i] = pts[3 * pc + 2];
            vx[i] = 0.0, vy[i] = 0.0, vz[i] = 0.0;
        }
        const unsigned k0 = chunk * pieces_per_chunk;
        const unsigned k1 = (k0 + pieces_per_chunk) < n_pieces ? (k0 + pieces_per_chunk) : n_pieces;
        for (unsigned k = k0 + w; k < k1; k += 4)
        {
            const ChainPiece pc = pieces[k];
            unsigned n = pc.n_seg;
            if (pc.row1)
            {
                const unsigned avail = cr - (pc.row1 - 1);
                n = n < avail ? n : avail;
            }
            const double sx = nodes[3 * pc.start_node + 0], sy = nodes[3 * pc.start_node + 1],
                         sz = nodes[3 * pc.start_node + 2];
            double ax[PPL], ay[PPL], az[PPL], aq[PPL]; // aq (|p - node|^2) only for the reference form
#pragma unroll
            for (int i = 0; i < PPL; ++i)
            {
                ax[i] = px[i] - sx, ay[i] = py[i] - sy, az[i] = pz[i] - sz;
                aq[i] = PREC == PVL_FP64_REF_FORM ? ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i] : 0.0;
            }
            const SegRec *sg = segs + pc.first_seg;
#pragma unroll 2
            for (unsigned j = 0; j < n; ++j)
            {
                const SegRec r = sg[j];
#pragma unroll
                for (int i = 0; i < PPL; ++i)
                {
                    const double bx = px[i] - r.nx, by = py[i] - r.ny, bz = pz[i] - r.nz;
                    bool cut;
                    if constexpr (PREC == PVL_FP64_REF_FORM)
                    {
                        const double bq = bx * bx + by * by + bz * bz;
                        const double m =
                            pair_m<PREC>(r.ex, r.ey, r.ez, r.len, ax[i], ay[i], az[i], aq[i], bx, by, bz, bq, tol, tol2, cut);
                        const double kx = cut ? 0.0 : (ay[i] * r.ez - az[i] * r.ey) * m;
                        const double ky = cut ? 0.0 : (az[i] * r.ex - ax[i] * r.ez) * m;
                        const double kz = cut ? 0.0 : (ax[i] * r.ey - ay[i] * r.ex) * m;
                        vx[i] = vx[i] + r.gamma * kx;
                        vy[i] = vy[i] + r.gamma * ky;
                        vz[i] = vz[i] + r.gamma * kz;
                        aq[i] = bq;
                    }
                    else
                    {
                        double cx, cy, cz;
                        const double m =
                            cross_m<PREC>(r.ex, r.ey, r.ez, r.len, ax[i], ay[i], az[i], tol2, cx, cy, cz, cut);
                        const double wm = cut ? 0.0 : r.gamma * m;
                        vx[i] = fma(wm, cx, vx[i]);
                        vy[i] = fma(wm, cy, vy[i]);
                        vz[i] = fma(wm, cz, vz[i]);
                    }
                    ax[i] = bx, ay[i] = by, az[i] = bz;
                }
            }
        }
#pragma unroll
        for (int i = 0; i < PPL; ++i)
        {
            red[w][0][i * kWave + lane] = vx[i];
            red[w][1][i * kWave + lane] = vy[i];
            red[w][2][i * kWave + lane] = vz[i];
        }
        __syncthreads();
        if (w == 0)
        {
#pragma unroll
            for (int i = 0; i < PPL; ++i)
            {
                const unsigned p = p0 + i * kWave;
                if (p >= n_pts)
                    continue;
                const unsigned q = i * kWave + lane;
                double sx2 = red[0][0][q], sy2 = red[0][1][q], sz2 = red[0][2][q];
                for (int u = 1; u < 4; ++u)
                {
                    sx2 = sx2 + red[u][0][q];
                    sy2 = sy2 + red[u][1][q];
                    sz2 = sz2 + red[u][2][q];
                }
                double *dst = partial + ((size_t)chunk * n_pts + p) * 3;
                dst[0] = sx2;
                dst[1] = sy2;
                dst[2] = sz2;
            }
        }
        __syncthreads(); // red is rewritten by the next item
    }
}

} // namespace

namespace pvl
{

// Point-velocity sum over the first n_active lines of the mesh (all lines for the public
// entry point; the wake passes the prefix of lines its filled rows use -- the other lines
// carry zero circulation).
// Chunking: aim for >= ~4 workgroups per CU while keeping >= 64 lines per wave.
static bool self_site(const char *prof_name) { return std::strcmp(prof_name, "wake_self_velocity") == 0; }
static unsigned site_points_per_group(const char *prof_name)
{
    return kWave * (self_site(prof_name) ? site_ppl<1>() : site_ppl<0>());
}

static void plan_chunks(unsigned n_pts, unsigned n_active, unsigned ppg, unsigned *n_chunks_out, unsigned *lpc_out)
{
    const unsigned pblocks = (n_pts + ppg - 1) / ppg;
    unsigned n_chunks = (2048 + pblocks - 1) / pblocks;
    const unsigned max_chunks = (n_active + 255) / 256;
    if (n_chunks > max_chunks)
        n_chunks = max_chunks;
    if (n_chunks < 1)
        n_chunks = 1;
    if (n_chunks > 65535u)
        n_chunks = 65535u;
    const unsigned lpc = (n_active + n_chunks - 1) / n_chunks;
    *n_chunks_out = (n_active + lpc - 1) / lpc;
    *lpc_out = lpc;
}

unsigned point_velocity_chunks(unsigned n_pts, unsigned n_active, const char *prof_name)
{
    unsigned nc, lpc;
    plan_chunks(n_pts, n_active, site_points_per_group(prof_name), &nc, &lpc);
    return nc;
}

size_t line_geom_bytes(unsigned n_lines) { return sizeof(LineGeom) * (size_t)(n_lines ? n_lines : 1); }

int point_velocity_partials(const pvl_mesh *m, double tol, const double *pos, const double *circ, unsigned n_pts,
                            const double *pts, int precision, unsigned n_active, const char *prof_name,
                            hipStream_t s, void *geom, double *gamma, double *partial, unsigned max_groups)
{
    LineGeom *g = (LineGeom *)geom;
    line_geom_kernel<<<(n_active + 255) / 256, 256, 0, s>>>(n_active, m->d_lines, pos, tol, g);
    PVL_HIP(hipGetLastError());
    line_circulation_kernel<<<(n_active + 255) / 256, 256, 0, s>>>(n_active, m->d_line_offsets, m->d_line_surfs,
                                                                     circ, gamma);
    PVL_HIP(hipGetLastError());
    unsigned n_chunks, lpc;
    const unsigned ppg = site_points_per_group(prof_name);
    plan_chunks(n_pts, n_active, ppg, &n_chunks, &lpc);
    const unsigned pblocks = (n_pts + ppg - 1) / ppg;
    const unsigned n_items = pblocks * n_chunks;
    const unsigned grid = max_groups && max_groups < n_items ? max_groups : n_items;
    ProfScope prof(prof_name, (double)n_pts * (double)n_active, s);
    dispatch_precision(precision, [&](auto P) {
        if (self_site(prof_name))
            point_velocity_kernel<decltype(P)::value, 1>
                <<<grid, kBlock, 0, s>>>(n_pts, pts, n_active, g, gamma, lpc, partial, tol, pblocks, n_items);
        else
            point_velocity_kernel<decltype(P)::value, 0>
                <<<grid, kBlock, 0, s>>>(n_pts, pts, n_active, g, gamma, lpc, partial, tol, pblocks, n_items);
    });
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

int line_circulations(const pvl_mesh *m, unsigned n_active, const double *circ, double *gamma, hipStream_t s)
{
    if (n_active == 0)
        return PVL_OK;
    line_circulation_kernel<<<(n_active + 255) / 256, 256, 0, s>>>(n_active, m->d_line_offsets, m->d_line_surfs,
                                                                     circ, gamma);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

int chain_segments(unsigned n_seg, const uint32_t *seg_ab, const uint32_t *seg_line, const double *nodes,
                   const double *gamma, unsigned n_active_lines, double tol, SegRec *out, hipStream_t s)
{
    if (n_seg == 0)
        return PVL_OK;
    chain_segment_kernel<<<(n_seg + 255) / 256, 256, 0, s>>>(n_seg, seg_ab, seg_line, nodes, gamma, n_active_lines,
                                                             tol, out);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

constexpr int kChainPPL = 2;

unsigned chain_velocity_chunks(unsigned n_pts, unsigned n_pieces)
{
    // aim at >= ~2048 work items while keeping >= 8 pieces (>= 2 per wave) per chunk
    const unsigned ppg = kWave * kChainPPL;
    const unsigned pblocks = (n_pts + ppg - 1) / ppg;
    unsigned n_chunks = (2048 + pblocks - 1) / pblocks;
    const unsigned max_chunks = (n_pieces + 7) / 8;
    if (n_chunks > max_chunks)
        n_chunks = max_chunks;
    if (n_chunks < 1)
        n_chunks = 1;
    const unsigned ppc = (n_pieces + n_chunks - 1) / n_chunks;
    return ppc ? (n_pieces + ppc - 1) / ppc : 1;
}

int chain_velocity_partials(unsigned n_pts, const double *pts, const double *nodes, const ChainPiece *pieces,
                            unsigned n_pieces, unsigned cr, const SegRec *segs, int precision, double tol,
                            double work, hipStream_t s, double *partial, unsigned max_groups)
{
    const unsigned n_chunks = chain_velocity_chunks(n_pts, n_pieces);
    const unsigned ppc = (n_pieces + n_chunks - 1) / n_chunks;
    const unsigned ppg = kWave * kChainPPL;
    const unsigned pblocks = (n_pts + ppg - 1) / ppg;
    const unsigned n_items = pblocks * n_chunks;
    const unsigned grid = max_groups && max_groups < n_items ? max_groups : n_items;
    ProfScope prof("wake_self_velocity", work, s);
    dispatch_precision(precision, [&](auto P) {
        chain_velocity_kernel<decltype(P)::value, kChainPPL><<<grid, kBlock, 0, s>>>(
            n_pts, pts, nodes, pieces, n_pieces, cr, segs, ppc, partial, tol, pblocks, n_items);
    });
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

int point_velocity_reduce(unsigned n_pts, unsigned n_chunks, const double *partial, double *out, int accumulate,
                          hipStream_t s)
{
    chunk_reduce_kernel<<<(3 * n_pts + 255) / 256, 256, 0, s>>>(n_pts, n_chunks, partial, out, accumulate);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

int point_velocity_impl(const pvl_mesh *m, double tol, const double *pos, const double *circ, unsigned n_pts,
                        const double *pts, double *out, int accumulate, int precision, unsigned n_active,
                        const char *prof_name)
{
    if (n_pts == 0)
        return PVL_OK;
    if (n_active > m->n_lines)
        n_active = m->n_lines;
    if (n_active == 0)
    {
        if (!accumulate)
            if (int st = device_fill32(out, 0u, sizeof(double) * 3 * n_pts))
                return st;
        return PVL_OK;
    }
    const unsigned n_chunks = point_velocity_chunks(n_pts, n_active, prof_name);
    void *g = scratch(0, line_geom_bytes(n_active));
    double *gamma = (double *)scratch(1, sizeof(double) * n_active);
    double *partial = (double *)scratch(2, sizeof(double) * 3 * (size_t)n_pts * n_chunks);
    if (!g || !gamma || !partial)
        return PVL_ERR_NOMEM;
    if (int st = point_velocity_partials(m, tol, pos, circ, n_pts, pts, precision, n_active, prof_name, stream(), g,
                                         gamma, partial, 0))
        return st;
    return point_velocity_reduce(n_pts, n_chunks, partial, out, accumulate, stream());
}

} // namespace pvl

extern "C" {

int pvl_induction_matrix3(const pvl_mesh *m, double tol, const double *pos, unsigned n_cpts, const double *cpts,
                          const double *normals, double *out, int precision)
{
    PVL_API_LOCK;
    if (int st = check_common(m, tol, precision))
        return st;
    return launch_surface<true>(m, tol, pos, n_cpts, cpts, normals, out, precision);
}

int pvl_induction_matrix(const pvl_mesh *m, double tol, const double *pos, unsigned n_cpts, const double *cpts,
                         double *out, int precision)
{
    PVL_API_LOCK;
    if (int st = check_common(m, tol, precision))
        return st;
    return launch_surface<false>(m, tol, pos, n_cpts, cpts, nullptr, out, precision);
}

int pvl_line_induction_matrix(const pvl_mesh *m, double tol, const double *pos, unsigned n_cpts, const double *cpts,
                              double *out, int precision)
{
    PVL_API_LOCK;
    if (int st = check_common(m, tol, precision))
        return st;
    if (n_cpts == 0 || m->n_lines == 0)
        return PVL_OK;
    LineGeom *g;
    if (int st = compute_line_geom(m, pos, tol, &g))
        return st;
    constexpr int lpw = 8;
    const dim3 grid((m->n_lines + 4 * lpw - 1) / (4 * lpw), (n_cpts + kCptTile - 1) / kCptTile);
    if (grid.y > 65535u)
    {
        set_error("too many control points (%u) for one launch", n_cpts);
        return PVL_ERR_ARG;
    }
    ProfScope prof("line_induction_matrix", (double)n_cpts * (double)m->n_lines);
    dispatch_precision(precision, [&](auto P) {
        line_induction_kernel<decltype(P)::value, lpw>
            <<<grid, kBlock, 0, stream()>>>(n_cpts, m->n_lines, g, cpts, out, tol);
    });
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

int pvl_point_velocity(const pvl_mesh *m, double tol, const double *pos, const double *circ, unsigned n_pts,
                       const double *pts, double *out, int accumulate, int precision)
{
    PVL_API_LOCK;
    if (int st = check_common(m, tol, precision))
        return st;
    return pvl::point_velocity_impl(m, tol, pos, circ, n_pts, pts, out, accumulate, precision, m->n_lines);
}

int pvl_surface_normal(const pvl_mesh *m, const double *pos, double *out)
{
    PVL_API_LOCK;
    if (int st = check_common(m, 0.0, PVL_FP64))
        return st;
    if (m->n_surfaces == 0)
        return PVL_OK;
    surface_geometry_kernel<<<(m->n_surfaces + 255) / 256, 256, 0, stream()>>>(
        m->n_surfaces, m->d_offsets, m->d_entries, m->d_lines, pos, out, nullptr);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

int pvl_surface_center(const pvl_mesh *m, const double *pos, double *out)
{
    PVL_API_LOCK;
    if (int st = check_common(m, 0.0, PVL_FP64))
        return st;
    if (m->n_surfaces == 0)
        return PVL_OK;
    surface_geometry_kernel<<<(m->n_surfaces + 255) / 256, 256, 0, stream()>>>(
        m->n_surfaces, m->d_offsets, m->d_entries, m->d_lines, pos, nullptr, out);
    PVL_HIP(hipGetLastError());
    return PVL_OK;
}

// Host-buffer forms: stage inputs, run, copy back, synchronise.
static int stage(const pvl_mesh *m, const double *pos, unsigned n_cpts, const double *cpts, const double *normals,
                 double **d_pos, double **d_cpts, double **d_norm)
{
    const size_t pb = sizeof(double) * 3 * (size_t)m->n_points, cb = sizeof(double) * 3 * (size_t)n_cpts;
    *d_pos = (double *)scratch(8, pb ? pb : 8);
    *d_cpts = (double *)scratch(9, cb ? cb : 8);
    *d_norm = normals ? (double *)scratch(10, cb ? cb : 8) : nullptr;
    if (!*d_pos || !*d_cpts || (normals && !*d_norm))
        return PVL_ERR_NOMEM;
    PVL_HIP(hipMemcpyAsync(*d_pos, pos, pb, hipMemcpyHostToDevice, stream()));
    PVL_HIP(hipMemcpyAsync(*d_cpts, cpts, cb, hipMemcpyHostToDevice, stream()));
    if (normals)
        PVL_HIP(hipMemcpyAsync(*d_norm, normals, cb, hipMemcpyHostToDevice, stream()));
    return PVL_OK;
}

static int unstage(double *out, const double *d_out, size_t bytes)
{
    PVL_HIP(hipMemcpyAsync(out, d_out, bytes, hipMemcpyDeviceToHost, stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    return PVL_OK;
}

int pvl_induction_matrix3_host(const pvl_mesh *m, double tol, const double *pos, unsigned n_cpts, const double *cpts,
                               const double *normals, double *out, int precision)
{
    PVL_API_LOCK;
    if (int st = check_common(m, tol, precision))
        return st;
    double *dp, *dc, *dn;
    if (int st = stage(m, pos, n_cpts, cpts, normals, &dp, &dc, &dn))
        return st;
    const size_t ob = sizeof(double) * (size_t)n_cpts * m->n_surfaces;
    double *d_out = (double *)scratch(11, ob ? ob : 8);
    if (!d_out)
        return PVL_ERR_NOMEM;
    if (int st = pvl_induction_matrix3(m, tol, dp, n_cpts, dc, dn, d_out, precision))
        return st;
    return unstage(out, d_out, ob);
}

int pvl_induction_matrix_host(const pvl_mesh *m, double tol, const double *pos, unsigned n_cpts, const double *cpts,
                              double *out, int precision)
{
    PVL_API_LOCK;
    if (int st = check_common(m, tol, precision))
        return st;
    double *dp, *dc, *dn;
    if (int st = stage(m, pos, n_cpts, cpts, nullptr, &dp, &dc, &dn))
        return st;
    const size_t ob = sizeof(double) * 3 * (size_t)n_cpts * m->n_surfaces;
    double *d_out = (double *)scratch(11, ob ? ob : 8);
    if (!d_out)
        return PVL_ERR_NOMEM;
    if (int st = pvl_induction_matrix(m, tol, dp, n_cpts, dc, d_out, precision))
        return st;
    return unstage(out, d_out, ob);
}

int pvl_line_induction_matrix_host(const pvl_mesh *m, double tol, const double *pos, unsigned n_cpts,
                                   const double *cpts, double *out, int precision)
{
    PVL_API_LOCK;
    if (int st = check_common(m, tol, precision))
        return st;
    double *dp, *dc, *dn;
    if (int st = stage(m, pos, n_cpts, cpts, nullptr, &dp, &dc, &dn))
        return st;
    const size_t ob = sizeof(double) * 3 * (size_t)n_cpts * m->n_lines;
    double *d_out = (double *)scratch(11, ob ? ob : 8);
    if (!d_out)
        return PVL_ERR_NOMEM;
    if (int st = pvl_line_induction_matrix(m, tol, dp, n_cpts, dc, d_out, precision))
        return st;
    return unstage(out, d_out, ob);
}

} // extern "C"
