// Internal declarations shared by the libpvl_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/pvl.h"

namespace pvl
{

constexpr uint32_t kInvalidId = 0x7FFFFFFFu; // src/core/common.h:54
constexpr uint32_t kReversed = 0x80000000u;  // src/core/common.h:55

// The library's scratch slots, LU/GMRES workspaces, streams and profiler tables are shared
// by every caller in the process.  Each entry point that touches them holds this lock from
// its first use of that state until it returns, so concurrent calls from several host
// threads are serialised and their device work is ordered on the library stream exactly
// as if one thread had issued it.  (The reference releases the GIL around its compute,
// src/meshobject.c:462-468, so callers may well be concurrent.)  Recursive: entry points
// call each other.
std::recursive_mutex &api_mutex();
#define PVL_API_LOCK std::lock_guard<std::recursive_mutex> _pvl_api_lock(::pvl::api_mutex())

// Errors raised by device work after its entry point returned (a single-launch triangular
// solve whose bounded wait timed out): kernels record them in pinned host memory; this
// returns PVL_ERR_HIP with a message (and clears the record) once the stream that ran them
// has been synchronised, PVL_OK otherwise.
int take_async_error();

// Launch of a kernel whose workgroups wait for each other (grid barriers), with grid size
// checked against what the device holds at once.  By default a cooperative launch; with
// PVL_COOPERATIVE=0 in the environment an ordinary launch of the same grid (every
// grid-wide wait in the library is bounded and reports an error rather than hanging), for
// tools that cannot follow cooperative launches.
int launch_grid_synced(const void *func, unsigned grid, unsigned block, void **args, size_t lds, hipStream_t s);

// Release of the process-lifetime workspaces (pvl_shutdown).
void lu_release_workspace();
void gmres_release_workspace();

// Error state (thread local) and the library stream.
void set_error(const char *fmt, ...);
hipStream_t stream();
hipStream_t side_stream(); // nullptr (error set) if it cannot be created
int side_groups();         // workgroup cap for grid-stride kernels on the side stream

// Kernels on the library stream's critical path raise their waves' issue priority above
// the side stream's (priority 0), which shares their SIMDs.
template <int LEVEL = 1> __device__ __forceinline__ void main_stream_priority() { __builtin_amdgcn_s_setprio(LEVEL); }
int ensure_init();

// Kernel timing hook: when profiling is on, brackets the launches made on stream s (default:
// the library stream) during the scope's lifetime with HIP events and accumulates the
// elapsed time under `name`.
struct ProfScope
{
    const char *name;
    double work;
    hipStream_t strm;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    explicit ProfScope(const char *n, double work_units = 0.0, hipStream_t s = nullptr);
    ~ProfScope();
};

int check_hip(hipError_t e, const char *what);
#define PVL_HIP(call)                                                                                                  \
    do                                                                                                                 \
    {                                                                                                                  \
        const int _st = ::pvl::check_hip((call), #call);                                                               \
        if (_st != PVL_OK)                                                                                             \
            return _st;                                                                                                \
    } while (0)

// Scratch device buffer reused across calls, grown on demand (one per slot).
void *scratch(int slot, size_t bytes);

// Device-to-device copy / 32-bit-word fill on stream s (nullptr: the library stream) with the
// library's own kernels (device_fill32 needs 4-byte aligned whole words).
int device_copy(void *dst, const void *src, size_t bytes, hipStream_t s = nullptr);
int device_fill32(void *dst, uint32_t value, size_t bytes, hipStream_t s = nullptr);

} // namespace pvl

struct pvl_mesh;
namespace pvl
{
// Point-velocity sum restricted to the first n_active lines of the mesh.
// prof_name labels the main kernel's launch for the event profiler; its work unit is the
// number of (line, point) pairs evaluated.
int point_velocity_impl(const pvl_mesh *m, double tol, const double *pos, const double *circ, unsigned n_pts,
                        const double *pts, double *out, int accumulate, int precision, unsigned n_active,
                        const char *prof_name = "point_velocity");

// The same sum in two parts with caller-owned buffers and stream.  point_velocity_chunks
// gives the line-chunk count (n_pts, n_active > 0; the chunking depends on the launch
// site prof_name, as the kernel's points per lane do); point_velocity_partials writes the
// per-chunk partial sums (geom: n_active LineGeom, gamma: n_active doubles, partial:
// 3 n_pts n_chunks doubles) on stream s with at most max_groups workgroups (0: one per work
// item); point_velocity_reduce adds them, chunk by chunk in order, to out (or to 0 when
// accumulate is 0) on stream s.
unsigned point_velocity_chunks(unsigned n_pts, unsigned n_active, const char *prof_name);
size_t line_geom_bytes(unsigned n_lines);
int point_velocity_partials(const pvl_mesh *m, double tol, const double *pos, const double *circ, unsigned n_pts,
                            const double *pts, int precision, unsigned n_active, const char *prof_name,
                            hipStream_t s, void *geom, double *gamma, double *partial, unsigned max_groups);
int point_velocity_reduce(unsigned n_pts, unsigned n_chunks, const double *partial, double *out, int accumulate,
                          hipStream_t s);

// Single-workgroup LU panel (lu_panel_f64.hip / lu_panel_f32.hip): factors the m x jb panel
// at P (row stride lda; matrix rows row_base..) with partial pivoting, piv[row_base + j]
// receiving the 0-based interchange row, status[0] the first zero pivot (1-based column).
constexpr int kPanelWgCols = 32;
constexpr int panel_wg_rows(size_t elem_size) { return elem_size == 8 ? 1024 : 2048; }
int panel_wg(double *P, int lda, int m, int jb, int row_base, int *piv, int *status);
int panel_wg(float *P, int lda, int m, int jb, int row_base, int *piv, int *status);
} // namespace pvl

// Line geometry precomputed once per call: start point, end point, unit direction and
// length, exactly as the reference's per-line prologue (src/core/flow_solver.c:95-103).
struct alignas(16) LineGeom
{
    double r1x, r1y, r1z;
    double r2x, r2y, r2z;
    double ex, ey, ez;
    double len;
};
static_assert(sizeof(LineGeom) == 80, "LineGeom layout");

// The device-resident mesh.
struct pvl_mesh
{
    unsigned n_points = 0, n_lines = 0, n_surfaces = 0;
    unsigned max_surface_lines = 0; // longest surface (entries)
    uint32_t *d_lines = nullptr;    // [2*n_lines] (bit 31 of each end kept as given)
    uint32_t *d_offsets = nullptr;  // [n_surfaces+1]
    uint32_t *d_entries = nullptr;  // [offsets[n_surfaces]]
    // line -> (surface, sign) transpose, for line circulations
    uint32_t *d_line_offsets = nullptr; // [n_lines+1]
    uint32_t *d_line_surfs = nullptr;   // [offsets[n_surfaces]] surface | kReversed
};
