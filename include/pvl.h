/*
 * pvl.h -- C ABI of libpvl_hip.so, the gfx950 implementation of PyVL's hot path:
 * line-vortex induction (matrix assembly and point-velocity sums), the dense
 * circulation solve, and the wake shed/convect step.
 *
 * Conventions shared by every entry point:
 *   - Plain pointers and sizes only.  "dev" pointers are device (HBM) buffers that the
 *     caller owns (allocate them with pvl_malloc or any HIP allocator); "host" pointers
 *     are ordinary host memory.
 *   - Vectors are AoS float64 triples, 24 bytes with no padding (the layout of the
 *     reference's real3_t, src/core/common.h:63-77).
 *   - Line connectivity is two uint32 per line; surface CSR is uint32 offsets[n+1] plus
 *     uint32 entries whose bit 31 is the orientation flag and bits 0..30 the line index
 *     (the reference's geo_id_t / line_t, src/core/common.h:46-56, src/core/mesh.h:10-14).
 *   - Every function returns an int status (PVL_OK = 0, negative on failure) and sets a
 *     thread-local message readable with pvl_last_error().
 *   - All device work is issued on one library stream per process (default priority),
 *     except the work pvl_wake_begin_update starts on a second library stream (also default
 *     priority, ordered with events); functions that return device results are
 *     asynchronous unless stated otherwise, and pvl_synchronize() waits for both streams.
 *   - Thread safety: entry points may be called from several host threads at once (the
 *     reference releases the GIL around its compute, src/meshobject.c:462-468).  The
 *     library's shared scratch, workspaces and streams are guarded by one process-wide
 *     lock, so such calls run one after another, in the order they take the lock; the
 *     host topology functions (pvl_mesh_from_elements, pvl_mesh_dual, pvl_wake_topology)
 *     touch no shared state and run concurrently.  Handles (pvl_mesh, pvl_lu, pvl_wake,
 *     pvl_mlp) must not be destroyed while another thread uses them.
 */
#ifndef PVL_H
#define PVL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pvl_status
{
    PVL_OK = 0,
    PVL_ERR_ARG = -1,      /* invalid argument (maps to ValueError/TypeError in the Python shim) */
    PVL_ERR_HIP = -2,      /* a HIP runtime call failed */
    PVL_ERR_NOMEM = -3,    /* device or host allocation failed (MemoryError) */
    PVL_ERR_SINGULAR = -4, /* exact zero pivot met in an LU factorisation */
    PVL_ERR_NOTCONV = -5,  /* iterative method did not reach its tolerance */
    PVL_ERR_NODEV = -6,    /* no usable gfx950 device */
};

enum pvl_precision
{
    /* FP64 throughout.  Dot products, h^2 and the cut-off test in the reference's operation
     * order without FMA; the subtended angle in the single-angle form
     * atan2((t2-t1) h, h^2 + t1 t2) with this library's atan2. */
    PVL_FP64 = 0,
    /* mixed: FP64 geometry, cut-off test, h^2 and cross product; FP32 angle (~1e-7). */
    PVL_FP32 = 1,
    /* FP64 with the reference's own angle expression atan2(t2,h) - atan2(t1,h) (two calls of
     * the HIP math library's atan2): the literal restatement, kept for cross-checks. */
    PVL_FP64_REF = 2,
};

/* ------------------------------------------------------------------------------------
 * Runtime
 * ---------------------------------------------------------------------------------- */
const char *pvl_last_error(void);
int pvl_init(int device);                 /* select device, create the library stream */
int pvl_device_count(int *count);
/* Waits for both library streams.  Also reports (PVL_ERR_HIP) an error that device work
 * recorded after its entry point had returned: a single-launch triangular solve whose
 * bounded wait timed out (its rows were written as NaN). */
int pvl_synchronize(void);
/* Releases the library's process-lifetime state (streams, events, pinned status words,
 * scratch and solver workspaces) after waiting for its streams.  Call it before the HIP
 * runtime is torn down (the Python shim registers it with atexit).  Handles created before
 * it stay valid for destruction; any later call re-initialises the library on device 0. */
int pvl_shutdown(void);
int pvl_malloc(void **dev_ptr, size_t bytes);
int pvl_free(void *dev_ptr);
int pvl_memcpy_h2d(void *dev_dst, const void *host_src, size_t bytes);
int pvl_memcpy_d2h(void *host_dst, const void *dev_src, size_t bytes);
int pvl_memcpy_d2d(void *dev_dst, const void *dev_src, size_t bytes);
int pvl_memset_zero(void *dev_dst, size_t bytes);

/* Kernel timing: when enabled, the main kernel of each entry point is bracketed by HIP
 * events on the library stream; totals are kept per entry-point name. */
int pvl_prof_enable(int on);
/* Restrict the timing to scopes named `name` (NULL: time every scope), so that a timed
 * region carries the events of one kernel only. */
int pvl_prof_filter(const char *name);
int pvl_prof_reset(void);
int pvl_prof_query(const char *name, double *total_ms, long *count);
/* Work units accumulated under `name` (for induction kernels: (line, point) pairs). */
int pvl_prof_work(const char *name, double *work);
/* Generic event timer on the library stream (for whole-step wall times). */
int pvl_timer_start(void);
int pvl_timer_stop(double *ms);

/* ------------------------------------------------------------------------------------
 * Host topology (replaces src/core/mesh.c:175-247 mesh_from_elements,
 * src/core/mesh.c:70-173 mesh_dual_from_primal, src/core/mesh.c:249-301 mesh_to_elements,
 * and src/meshobject.c:780-860 merge).  Same line numbering (first appearance order) and
 * orientation bits, built with a hash map in O(total points) instead of O(E*L).
 * ---------------------------------------------------------------------------------- */
/* lines_out must hold 2*sum(point_counts) uint32, surface_offsets_out n_elements+1,
 * surface_lines_out sum(point_counts). */
int pvl_mesh_from_elements(unsigned n_elements, const unsigned *point_counts, const unsigned *flat_points,
                           unsigned *n_lines_out, uint32_t *lines_out, uint32_t *surface_offsets_out,
                           uint32_t *surface_lines_out);
/* Dual: dual_lines_out[2*n_lines] (INVALID_ID = 0x7FFFFFFF where absent),
 * dual_offsets_out[n_points+1], dual_entries_out[2*n_lines]. */
int pvl_mesh_dual(unsigned n_points, unsigned n_lines, const uint32_t *lines, unsigned n_surfaces,
                  const uint32_t *surface_offsets, const uint32_t *surface_lines, uint32_t *dual_lines_out,
                  uint32_t *dual_offsets_out, uint32_t *dual_entries_out);
/* Wake ring mesh of python/pyvl/wake_models.py:35-52 built in closed form:
 * n_strips strips of line_rows rings, nodes 2*(line_rows+1)*n_strips.  Output sizes:
 * lines_out 2*n_lines with n_lines = n_strips*(3*line_rows+1), offsets n_strips*line_rows+1,
 * entries 4*n_strips*line_rows. */
int pvl_wake_topology(unsigned n_strips, unsigned line_rows, unsigned *n_lines_out, uint32_t *lines_out,
                      uint32_t *surface_offsets_out, uint32_t *surface_lines_out);

/* ------------------------------------------------------------------------------------
 * Device mesh: the topology uploaded once, plus the line->surface transpose used by the
 * point-velocity sums.  Replaces the mesh_t owned by cvl.Mesh (src/meshobject.c:25-124).
 * ---------------------------------------------------------------------------------- */
typedef struct pvl_mesh pvl_mesh;
int pvl_mesh_create(pvl_mesh **out, unsigned n_points, unsigned n_lines, const uint32_t *lines, unsigned n_surfaces,
                    const uint32_t *surface_offsets, const uint32_t *surface_lines);
int pvl_mesh_destroy(pvl_mesh *mesh);
int pvl_mesh_counts(const pvl_mesh *mesh, unsigned *n_points, unsigned *n_lines, unsigned *n_surfaces);

/* ------------------------------------------------------------------------------------
 * Induction (device buffers).
 * ---------------------------------------------------------------------------------- */
/* Mesh.induction_matrix3 (src/meshobject.c:384-473 = src/core/flow_solver.c:86-139 then
 * :174-205): out[n_cpts][n_surfaces] = (sum over surface lines of +-K(line, cpt)) . normal.
 * Fused: no line buffer. */
int pvl_induction_matrix3(const pvl_mesh *mesh, double tol, const double *positions_dev, unsigned n_cpts,
                          const double *cpts_dev, const double *normals_dev, double *out_dev, int precision);
/* Mesh.induction_matrix (src/meshobject.c:475-556 = flow_solver.c:86-172):
 * out[n_cpts][n_surfaces][3]. */
int pvl_induction_matrix(const pvl_mesh *mesh, double tol, const double *positions_dev, unsigned n_cpts,
                         const double *cpts_dev, double *out_dev, int precision);
/* Mesh.line_induction_matrix (src/meshobject.c:1232-1286 = flow_solver.c:86-139):
 * out[n_cpts][n_lines][3]. */
int pvl_line_induction_matrix(const pvl_mesh *mesh, double tol, const double *positions_dev, unsigned n_cpts,
                              const double *cpts_dev, double *out_dev, int precision);
/* Point-velocity sum: the reference materialises induction_matrix and then takes
 * np.vecdot(ind, circulation[None,:,None], axis=1) (python/pyvl/wake_models.py:125-126,
 * :170-175, python/pyvl/solver.py:274-280).  Here out[i] (+)= sum_surface Gamma_s * K_s(p_i)
 * is evaluated as sum over lines of gamma_line * K(line, p_i) with no tensor in HBM.
 * accumulate != 0 adds into out instead of overwriting. */
int pvl_point_velocity(const pvl_mesh *mesh, double tol, const double *positions_dev, const double *circulation_dev,
                       unsigned n_pts, const double *pts_dev, double *out_dev, int accumulate, int precision);
/* No counterpart in the reference.  pvl_point_velocity for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): the same kernels and chunking, so bit-identical results, but its working
 * buffers (line geometry and line circulations, 88 B per line, plus per-chunk partial sums,
 * 24 B per point and line chunk) live in one fixed array of PVL_NOALLOC_SCRATCH_BYTES per
 * GPU, which the library holds until the process ends; a call that needs more fails with
 * PVL_ERR_ARG before any launch.  pvl_init must have been called (it looks up that array's
 * device address).  Every call on a GPU shares the array: calls whose device work overlaps
 * (e.g. captured replays on other streams) race.  Allocates nothing while the event
 * profiler (pvl_prof_enable) is off. */
#define PVL_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int pvl_point_velocity_noalloc(const pvl_mesh *mesh, double tol, const double *positions_dev,
                               const double *circulation_dev, unsigned n_pts, const double *pts_dev, double *out_dev,
                               int accumulate, int precision);
/* Mesh.surface_normal / Mesh.surface_average_vec3 (src/meshobject.c:956-1062 =
 * src/core/mesh.c:48-68 and :24-46). */
int pvl_surface_normal(const pvl_mesh *mesh, const double *positions_dev, double *out_dev);
int pvl_surface_center(const pvl_mesh *mesh, const double *positions_dev, double *out_dev);
/* Mesh.line_forces (src/meshobject.c:1288-1396), the force on every primal line from its
 * reduced line circulation: for line (p1, p2) whose dual line is (s+, s-),
 * c = circulation[s+] - circulation[s-] (INVALID_ID ends skipped) and
 * out[l] = (r_p2 - r_p1) x ((v_p1 + v_p2) * (0.5 c)).  dual_lines_dev holds the dual mesh's
 * lines (2 uint32 per primal line, surface indices or INVALID_ID; an index at or beyond the
 * primal's surface count counts as absent), the dual's counts are checked against the
 * primal's transposed (points <-> surfaces) as the reference does.  circulation_dev[primal
 * n_surfaces], positions_dev / velocity_dev [primal n_points][3], out_dev [n_lines][3]. */
int pvl_line_forces(const pvl_mesh *primal, const uint32_t *dual_lines_dev, unsigned dual_n_points,
                    unsigned dual_n_lines, unsigned dual_n_surfaces, const double *circulation_dev,
                    const double *positions_dev, const double *velocity_dev, double *out_dev);

/* Host-buffer convenience forms of the three tensor entry points: stage through device
 * memory and return when the result is in out_host.  These are what a C caller such as
 * the reference's src/meshobject.c would call in place of compute_line_induction +
 * line_induction_to_(normal_)surface_induction. */
int pvl_induction_matrix3_host(const pvl_mesh *mesh, double tol, const double *positions, unsigned n_cpts,
                               const double *cpts, const double *normals, double *out, int precision);
int pvl_induction_matrix_host(const pvl_mesh *mesh, double tol, const double *positions, unsigned n_cpts,
                              const double *cpts, double *out, int precision);
int pvl_line_induction_matrix_host(const pvl_mesh *mesh, double tol, const double *positions, unsigned n_cpts,
                                   const double *cpts, double *out, int precision);

/* ------------------------------------------------------------------------------------
 * Dense solve.  Replaces scipy.linalg.lu_factor / lu_solve as called by run_solver
 * (python/pyvl/solver.py:223-227; LAPACK dgetrf/dgetrs semantics: P A = L U with partial
 * pivoting, piv[j] = row interchanged with row j, 0-based).  Matrices are row-major
 * n x n on the device; right-hand sides are n x nrhs row-major.
 * ---------------------------------------------------------------------------------- */
typedef struct pvl_lu pvl_lu;
/* precision PVL_FP64: factors kept in FP64; PVL_FP32: factors kept in FP32 (MFMA FP32). */
int pvl_lu_create(pvl_lu **out, int n, int precision);
int pvl_lu_destroy(pvl_lu *h);
/* Pivoting of the factorisation:
 *   PVL_PIVOT_PARTIAL (default): partial pivoting, LAPACK's interchanges;
 *   PVL_PIVOT_AUTO: the same factorisation, computed without interchanges when partial
 *     pivoting would make none (checked at every column's elimination: no |a_ij| above
 *     |a_jj|), else refactored with partial pivoting -- the case of panel-method systems
 *     whose diagonal dominates (thin lifting surfaces);
 *   PVL_PIVOT_NONE: no interchanges at all (an FP32 factorisation for FP64 refinement on
 *     diagonally strong systems).  No reference counterpart for AUTO and NONE. */
enum pvl_pivoting
{
    PVL_PIVOT_NONE = 0,
    PVL_PIVOT_PARTIAL = 1,
    PVL_PIVOT_AUTO = 2
};
int pvl_lu_set_pivoting(pvl_lu *h, int pivoting);
/* Copy A (FP64, device) into the handle (converting for FP32) and factor it.  Returns
 * PVL_ERR_SINGULAR with *info_out = first zero pivot (1-based) when U is exactly singular. */
int pvl_lu_factor(pvl_lu *h, const double *A_dev, int *info_out);
/* Solve in the handle's precision: B_dev (double* for FP64 handles, float* for FP32) is
 * overwritten with the solution.  1 <= nrhs <= 32.  Asynchronous on the library stream;
 * solves of one handle are ordered by that stream (they share its flag and staging
 * buffers).  Large systems run as one launch per triangle whose workgroups wait on each
 * other with bounded waits: should a wait time out (about a second), the affected rows are
 * written as NaN rather than left stale. */
int pvl_lu_solve(pvl_lu *h, void *B_dev, int nrhs);
int pvl_lu_pivots(const pvl_lu *h, int *piv_host);
int pvl_lu_factors(const pvl_lu *h, void *lu_dev_out);
/* R = B - A X in FP64. */
int pvl_residual(int n, const double *A_dev, const double *X_dev, const double *B_dev, double *R_dev, int nrhs);
/* Mixed precision solve: h holds the FP32 factorisation of A (pvl_lu_factor on an FP32
 * handle); X_dev receives the FP64 solution of A X = B after FP64 iterative refinement from
 * X = 0, stopping when every column has max|b - A x| <= tol * max|b| (FP64 residuals).
 * iters_out: corrections; res_out[nrhs]: final relative residuals.  PVL_ERR_NOTCONV when
 * the residual stalls or max_iter corrections do not suffice.  No reference counterpart:
 * the reference only has the FP64 LAPACK solve. */
int pvl_solve_refine(pvl_lu *h, const double *A_dev, const double *B_dev, double *X_dev, int nrhs, double tol,
                     int max_iter, int *iters_out, double *res_out);
/* Closed-body deflation A[i][j] += c on the body's diagonal block [r0, r1): the rank-one
 * shift that makes the closed-surface system (A 1 = 0, python/pyvl/solver.py:229-233)
 * nonsingular without changing its mean-free solution. */
int pvl_deflate(int n, double *A_dev, int r0, int r1, double c);

/* ------------------------------------------------------------------------------------
 * Unsteady line wake (python/pyvl/wake_models.py WakeModelLineExplicitUnsteady), resident
 * in HBM.  Coincident nodes of neighbouring strips are stored once; the reference layout is
 * available through pvl_wake_get_state / pvl_wake_set_state.
 * ---------------------------------------------------------------------------------- */
typedef struct pvl_wake pvl_wake;
/* WakeModelLineExplicitUnsteady.__init__ (:54-81): bordering_nodes and adjacent_surfaces
 * are (n_shed, 2) uint32 (INVALID_ID allowed in adjacent_surfaces). */
int pvl_wake_create(pvl_wake **out, unsigned n_shed, unsigned line_rows, const uint32_t *bordering_nodes,
                    const uint32_t *adjacent_surfaces, double vortex_tol, double time);
int pvl_wake_destroy(pvl_wake *w);
int pvl_wake_info(const pvl_wake *w, unsigned *n_te, unsigned *n_lines, unsigned *step_count, double *current_time);
/* update (:83-157) with a uniform free stream vinf[3].  mode 0: the reference's convection
 * (free stream + bound-mesh induction); mode 1: free wake (also the wake's own induction,
 * i.e. get_velocity of :159-175). */
int pvl_wake_update(pvl_wake *w, double time, const pvl_mesh *bound, const double *bound_pos_dev,
                    const double *circulation_dev, const double *vinf, int mode, int precision);
/* The same update with any free stream the reference supports (flow.get_velocity(time, nodes)
 * at :124): flow is the 7-double descriptor {kind, a[3], b[3]}: kind 0 FlowConditionsUniform
 * (velocity a), kind 1 FlowConditionsRotating (v = (p - a) x b, python/pyvl/flow_conditions.py
 * :332-446), evaluated at each node's position before the step. */
int pvl_wake_update_flow(pvl_wake *w, double time, const pvl_mesh *bound, const double *bound_pos_dev,
                         const double *circulation_dev, const double *flow, int mode, int precision);
/* Scheduling only, no reference counterpart: starts the wake's own induction of the next
 * pvl_wake_update (mode 1; it needs only the wake's current state) on a second library
 * stream, so that it runs beside the bound-circulation solve issued in between.  That
 * pvl_wake_update (same mode and precision, no pvl_wake_set_state in between) then adds the
 * partial sums in the same order: results are bit-identical to pvl_wake_update alone.  A
 * no-op for mode 0, an empty wake, or when already started. */
int pvl_wake_begin_update(pvl_wake *w, int mode, int precision);
/* apply_corrections (:177-223) on a row-major n_cpts x n_surf matrix and n_cpts rhs. */
int pvl_wake_apply_corrections(pvl_wake *w, unsigned n_cpts, const double *cpts_dev, const double *normals_dev,
                               double *mat_dev, unsigned n_surf, double *rhs_dev, int precision);
/* get_velocity (:159-175): out (n_pts x 3) (+)= wake-induced velocity. */
int pvl_wake_get_velocity(pvl_wake *w, unsigned n_pts, const double *pts_dev, double *out_dev, int accumulate,
                          int precision);
/* Reference layout: positions (n_shed, 2, R+1, 3), circulation (n_shed, R), host buffers. */
int pvl_wake_get_state(pvl_wake *w, double *positions_host, double *circulation_host);
int pvl_wake_set_state(pvl_wake *w, const double *positions_host, const double *circulation_host, unsigned step_count,
                       double current_time);

/* ------------------------------------------------------------------------------------
 * Learned warm start.  No reference counterpart (the reference solves every step with
 * LAPACK).  A small MLP (BF16 operands on MFMA, FP32 accumulation, FP32 master weights and
 * Adam state) predicts per-panel circulations; FP32 GMRES inside FP64 refinement starts from
 * that prediction; LU is the caller's fallback.
 * ---------------------------------------------------------------------------------- */
typedef struct pvl_mlp pvl_mlp;
/* n_layers linear layers: d_in -> width -> ... -> width -> 1, SiLU between layers. */
int pvl_mlp_create(pvl_mlp **out, int d_in, int width, int n_layers);
int pvl_mlp_destroy(pvl_mlp *h);
int pvl_mlp_num_params(const pvl_mlp *h, long *n);
/* Flat FP32 parameters, per layer W (out x in, row-major) then b (out). */
int pvl_mlp_set_params(pvl_mlp *h, const float *params_host);
int pvl_mlp_get_params(const pvl_mlp *h, float *params_host);
/* Y[b] = net(X[b, :]); X (batch x d_in) and Y (batch) FP32 device buffers. */
int pvl_mlp_forward(pvl_mlp *h, const float *X_dev, float *Y_dev, int batch);
/* Gradient of the mean squared error against T (no update). */
int pvl_mlp_gradient(pvl_mlp *h, const float *X_dev, const float *T_dev, int batch, float *grads_host,
                     float *loss_host);
/* One Adam step (L2 weight decay wd * p added to the gradient); loss_host may be NULL. */
int pvl_mlp_train_step(pvl_mlp *h, const float *X_dev, const float *T_dev, int batch, float lr, float beta1,
                       float beta2, float eps, float wd, float *loss_host);
/* Iterative solve of A x = b from the start x_dev: FP64 residuals r = b - A x, stop when
 * max|r| <= tol * max|b|, else one FP32 GMRES(m) cycle (right Jacobi preconditioning) on
 * A d = r, x += d, at most max_outer corrections.  stats[0] = corrections, stats[1] = GMRES
 * steps, stats[2] = converged flag; res_out[0] / res_out[1] = max|r|/max|b| at the start /
 * at the end.  PVL_ERR_NOTCONV when not converged (the caller falls back to LU).  No
 * reference counterpart (the reference factors every step, python/pyvl/solver.py:223-227). */
int pvl_gmres_refine(int n, const double *A64_dev, const float *A32_dev, const double *b_dev, double *x_dev, double tol,
                     int m, double inner_tol, int max_outer, int *stats, double *res_out);
/* FP32 copy of an FP64 device array. */
int pvl_convert_f64_f32(const double *src_dev, float *dst_dev, size_t count);

/* ------------------------------------------------------------------------------------
 * Step helpers of run_solver (python/pyvl/solver.py:199-206, :229-233).
 * ---------------------------------------------------------------------------------- */
/* out_i = n_i . -(vinf - cpvel_i); cpvel_dev may be NULL (fixed body). */
int pvl_rhs_normal_velocity(unsigned n, const double *normals_dev, const double *cpvel_dev, const double *vinf,
                            double *out_dev);
/* out_i = n_i . -(V(cp_i) - cpvel_i) for the free-stream descriptor of pvl_wake_update_flow
 * (cpts_dev may be NULL for a uniform stream). */
int pvl_rhs_flow(unsigned n, const double *normals_dev, const double *cpvel_dev, const double *cpts_dev,
                 const double *flow, double *out_dev);
/* x[r0:r1] -= mean(x[r0:r1]) (closed-body circulation gauge). */
int pvl_remove_mean(double *x_dev, int r0, int r1);

/* ------------------------------------------------------------------------------------
 * Rigid-frame kinematics (python/pyvl/solver.py:180-193; ReferenceFrame transforms
 * src/referenceframeobject.c:367-450, add_velocity :939-955 and its subclasses'
 * python/pyvl/reference_frames.py:122-147, :298-322).
 * ---------------------------------------------------------------------------------- */
enum pvl_frame_kind
{
    PVL_FRAME_NONE = 0,        /* ReferenceFrame: adds no velocity */
    PVL_FRAME_TRANSLATING = 1, /* v += c */
    PVL_FRAME_ROTOR = 2,       /* v += p x c (c = angular velocity) */
    PVL_FRAME_AFFINE = 3,      /* v += W p + c (oscillating frames) */
};
/* The time step's chain walk for every point of one body, child frame first: per frame
 * p = R^T (p - d); v += term(p); v = R^T v, starting from p = p0, v = 0.  terms holds 24
 * doubles per frame, already evaluated at the step's time: R^T row-major [9], d [3], W
 * row-major [9], c [3]; kinds one pvl_frame_kind per frame; at most 8 frames.  vel_dev
 * may be NULL (positions only). */
int pvl_frame_kinematics(unsigned n_points, const double *p0_dev, unsigned n_frames, const double *terms,
                         const int *kinds, double *pos_dev, double *vel_dev);

#ifdef __cplusplus
}
#endif

#endif /* PVL_H */
