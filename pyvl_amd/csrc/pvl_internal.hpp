// Internal declarations shared by the libpvl_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

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

// Launch of a kernel whose workgroups wait for each other (grid barriers, exchange rounds), with grid size
// checked against what the device holds at once.  By default an ordinary launch (every
// grid-wide wait in the library is bounded and reports an error rather than hanging); with
// PVL_COOPERATIVE=1 in the environment a cooperative launch of the same grid.
int launch_grid_synced(const void *func, unsigned grid, unsigned block, void **args, size_t lds, hipStream_t s);

// Release of the process-lifetime workspaces (pvl_shutdown).
void lu_release_workspace();
void gmres_release_workspace();

// Error state (thread local) and the library stream.
void set_error(const char *fmt, ...);
hipStream_t stream();
hipStream_t side_stream(); // nullptr (error set) if it cannot be created
int side_groups(int per_cu = 5); // workgroup cap for grid-stride kernels on the side stream

// Kernels on the library stream's critical path raise their waves' issue priority above
// the side stream's (priority 0), which shares their SIMDs.
template <int LEVEL = 1> __device__ __forceinline__ void main_stream_priority() { __builtin_amdgcn_s_setprio(LEVEL); }
int ensure_init();
// The device the library is bound to (pvl_init), or -1 before initialisation.
int initialised_device();
// Set-up of the no-allocation entry point on `device` (current): looks up and keeps the
// device address of its fixed scratch array.  Called once per device by initialisation.
int noalloc_setup(int device);

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

// Point-velocity sum over segment chains (the wake's self-induction).  A chain is a polyline
// of nodes; a piece is a run of up to kChainPiece consecutive segments of one chain, starting
// at node start_node, whose segment records are segs[first_seg .. first_seg + n_seg).  A
// piece with row1 > 0 runs streamwise from node row row1 - 1 and only its first cr - (row1 - 1)
// segments are active at cr filled rows.  Walking a piece, each node's p - r (and, in the
// reference form, |p - r|^2) is formed once and serves both segments that meet there.
constexpr unsigned kChainPiece = 32;
struct ChainPiece
{
    uint32_t start_node, first_seg, n_seg, row1;
};
// Per-step segment record: end node, unit direction and length (the reference's line
// prologue, src/core/flow_solver.c:95-103, in the chain's direction) and circulation.
struct alignas(16) SegRec
{
    double nx, ny, nz;
    double ex, ey, ez;
    double len, gamma;
};
static_assert(sizeof(SegRec) == 64, "SegRec layout");
// Line circulations gamma_l = sum of +-Gamma over the surfaces holding line l, lines < n_active.
int line_circulations(const pvl_mesh *m, unsigned n_active, const double *circ, double *gamma, hipStream_t s);
// SegRec of n_seg segments (seg_ab: node pairs; seg_line: mesh line | bit 31 when the chain runs
// against the line's direction); lines >= n_active_lines carry no circulation.
int chain_segments(unsigned n_seg, const uint32_t *seg_ab, const uint32_t *seg_line, const double *nodes,
                   const double *gamma, unsigned n_active_lines, double tol, SegRec *out, hipStream_t s);
unsigned chain_velocity_chunks(unsigned n_pts, unsigned n_pieces);
// Partial sums per piece chunk (3 n_pts doubles each, in chunk order; point_velocity_reduce adds
// them); work = pairs evaluated, for the profiler.
int chain_velocity_partials(unsigned n_pts, const double *pts, const double *nodes, const ChainPiece *pieces,
                            unsigned n_pieces, unsigned cr, const SegRec *segs, int precision, double tol,
                            double work, hipStream_t s, double *partial, unsigned max_groups);

// Single-workgroup LU panel (lu_panel_f64.hip / lu_panel_f32.hip): factors the m x jb panel
// at P (row stride lda; matrix rows row_base..) with partial pivoting, piv[row_base + j]
// receiving the 0-based interchange row, status[0] the first zero pivot (1-based column).
constexpr int kPanelWgCols = 32;
constexpr int panel_wg_rows(size_t elem_size) { return elem_size == 8 ? 1024 : 2048; }
int panel_wg(double *P, int lda, int m, int jb, int row_base, int *piv, int *status);
int panel_wg(float *P, int lda, int m, int jb, int row_base, int *piv, int *status);
} // namespace pvl

// Free-stream velocity field on the device (python/pyvl/flow_conditions.py): uniform
// (kind 0, velocity a) or rotating about a point (kind 1, v = (p - a) x b,
// FlowConditionsRotating.get_velocity :412-446 with numpy's cross product order).
struct FlowField
{
    int kind;
    double a[3], b[3];
};
// From the C ABI's 7-double descriptor {kind, a[3], b[3]}; PVL_ERR_ARG on an unknown kind.
int flow_from_descriptor(const double *desc, FlowField *out);
__device__ __forceinline__ void flow_velocity(const FlowField &f, double x, double y, double z, double &vx,
                                              double &vy, double &vz)
{
    if (f.kind == 0)
    {
        vx = f.a[0], vy = f.a[1], vz = f.a[2];
        return;
    }
    const double px = x - f.a[0], py = y - f.a[1], pz = z - f.a[2];
    vx = py * f.b[2] - pz * f.b[1];
    vy = pz * f.b[0] - px * f.b[2];
    vz = px * f.b[1] - py * f.b[0];
}

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
    // assembly program (pvl::build_assembly_program): per run of up to kAsmGroups groups
    uint32_t *d_asm_off = nullptr; // [n_runs + 1] entry offsets
    uint32_t *d_asm_grp = nullptr; // [n_runs * kAsmGroups] groups of the run (kAsmNoGroup: none)
    uint32_t *d_asm = nullptr;     // [2 * entries] (line | kReversed, op)
    unsigned n_asm_runs = 0;
};

namespace pvl
{
// The assembly (induction_matrix3) walks its surfaces in groups of kAsmSurfaces consecutive
// surfaces (group g: surfaces 8g .. 8g + 7) and its groups in runs of up to kAsmGroups, one
// run per wave; a run's groups are chosen by shared lines (on a structured body: the same
// columns of kAsmGroups neighbouring rows), so that most lines of a run serve two of its
// surfaces.  Inside a run each line's n . K is evaluated once and kept in one of kAsmSlots
// per-lane LDS slots while a later surface of the run still needs it.  The program lists,
// per run, group by group and in CSR order, one entry per surface line: word 0 the line with
// the orientation bit, word 1 the operation (kAsmEval / kAsmEvalStore / kAsmLoad / kAsmNone,
// the slot in bits 2-5, kAsmEnd on a surface's last entry; a surface without lines gets one
// kAsmNone entry with kAsmEnd).  A group with an odd entry count is padded with one kAsmNone
// entry without kAsmEnd (adds zero), so that groups start on even entries (the kernel takes
// entries in pairs).  Built on the host once per mesh, in O(entries).
constexpr unsigned kAsmSurfaces = 8;
constexpr unsigned kAsmGroups = 4;
constexpr unsigned kAsmSlots = 10;
constexpr uint32_t kAsmEval = 0, kAsmEvalStore = 1, kAsmLoad = 2, kAsmNone = 3, kAsmEnd = 1u << 8;
constexpr uint32_t kAsmNoGroup = 0xFFFFFFFFu;
void build_assembly_program(unsigned n_surfaces, const uint32_t *offsets, const uint32_t *entries,
                            std::vector<uint32_t> &run_off, std::vector<uint32_t> &run_grp,
                            std::vector<uint32_t> &prog);
} // namespace pvl
