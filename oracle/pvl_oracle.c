/*
 * ORACLE -- test infrastructure only.  A plain-C FP64 restatement of PyVL's hot path used
 * as the checker for libpvl_hip.so and as the CPU baseline's "port" leg.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product never
 * does.
 *
 * Built like the reference (oracle/Makefile): gcc -O3, no -march (x86-64 baseline: no FMA
 * instructions), -ffp-contract=off, OpenMP.  Every function restates one reference
 * function, cited by file:line under /root/reference, keeping its arithmetic in the same
 * order so that the results are bit-identical to the reference's C on the same machine
 * (pinned by tests/test_oracle.py against oracle/_ref, the reference's own sources compiled
 * unmodified).
 *
 * Layouts: vectors are AoS double triples; lines are uint32 pairs; surface entries are
 * uint32 with bit 31 = reversed and bits 0..30 = line index (src/core/common.h:46-56).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MASK 0x7FFFFFFFu
#define ORC_REV 0x80000000u
#define ORC_INVALID 0x7FFFFFFFu

/* K(line, p) exactly as src/core/flow_solver.c:95-136 (line prologue :95-103, pair
 * :114-135). */
typedef struct
{
    double r1[3], r2[3], e[3], len;
} orc_line;

static orc_line orc_line_geom(const uint32_t *lines, const double *pos, unsigned il)
{
    orc_line g;
    const uint32_t a = lines[2 * il] & ORC_MASK, b = lines[2 * il + 1] & ORC_MASK;
    for (int c = 0; c < 3; ++c)
    {
        g.r1[c] = pos[3 * a + c];
        g.r2[c] = pos[3 * b + c];
    }
    const double d0 = g.r2[0] - g.r1[0], d1 = g.r2[1] - g.r1[1], d2 = g.r2[2] - g.r1[2];
    g.len = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    g.e[0] = d0 / g.len;
    g.e[1] = d1 / g.len;
    g.e[2] = d2 / g.len;
    return g;
}

static void orc_pair(const orc_line *g, const double *p, double tol, double *k)
{
    if (g->len < tol)
    {
        k[0] = k[1] = k[2] = 0.0;
        return;
    }
    const double a0 = p[0] - g->r1[0], a1 = p[1] - g->r1[1], a2 = p[2] - g->r1[2];
    const double b0 = p[0] - g->r2[0], b1 = p[1] - g->r2[1], b2 = p[2] - g->r2[2];
    const double t1 = g->e[0] * a0 + g->e[1] * a1 + g->e[2] * a2;
    const double t2 = g->e[0] * b0 + g->e[1] * b1 + g->e[2] * b2;
    const double n1 = (a0 * a0 + a1 * a1 + a2 * a2) - (t1 * t1);
    const double n2 = (b0 * b0 + b1 * b1 + b2 * b2) - (t2 * t2);
    const double h2 = (n1 + n2) / 2.0;
    if (h2 < tol * tol)
    {
        k[0] = k[1] = k[2] = 0.0;
        return;
    }
    const double h = sqrt(h2);
    const double m = (atan2(t2, h) - atan2(t1, h)) / h;
    /* cross(dr1, e) scaled by m */
    k[0] = (a1 * g->e[2] - a2 * g->e[1]) * m;
    k[1] = (a2 * g->e[0] - a0 * g->e[2]) * m;
    k[2] = (a0 * g->e[1] - a1 * g->e[0]) * m;
}

/* compute_line_induction (src/core/flow_solver.c:86-139): out[(icp*n_lines+il)*3+c]. */
void orc_line_induction(unsigned n_lines, const uint32_t *lines, const double *pos, unsigned n_cpts,
                        const double *cpts, double *out, double tol)
{
    orc_line *geo = malloc(sizeof(*geo) * (n_lines ? n_lines : 1));
    for (unsigned il = 0; il < n_lines; ++il)
        geo[il] = orc_line_geom(lines, pos, il);
    long icp;
#pragma omp parallel for schedule(static)
    for (icp = 0; icp < (long)n_cpts; ++icp)
        for (unsigned il = 0; il < n_lines; ++il)
            orc_pair(geo + il, cpts + 3 * icp, tol, out + ((size_t)icp * n_lines + il) * 3);
    free(geo);
}

/* Surface reduction (src/core/flow_solver.c:149-171 and :182-204) for one control point
 * given its line buffer row `krow` (n_lines*3). */
static void orc_reduce(unsigned s, const uint32_t *off, const uint32_t *ent, const double *krow, double *res)
{
    res[0] = res[1] = res[2] = 0.0;
    for (uint32_t k = off[s]; k < off[s + 1]; ++k)
    {
        const uint32_t e = ent[k];
        const double *v = krow + 3 * (size_t)(e & ORC_MASK);
        if (e & ORC_REV)
        {
            res[0] = res[0] - v[0];
            res[1] = res[1] - v[1];
            res[2] = res[2] - v[2];
        }
        else
        {
            res[0] = res[0] + v[0];
            res[1] = res[1] + v[1];
            res[2] = res[2] + v[2];
        }
    }
}

/* Mesh.induction_matrix3 / induction_matrix (src/meshobject.c:384-556): the line buffer is
 * built one control-point row at a time (same values as the full buffer) and reduced. */
static void orc_surface_common(unsigned n_lines, const uint32_t *lines, unsigned n_surf, const uint32_t *off,
                               const uint32_t *ent, const double *pos, unsigned n_cpts, const double *cpts,
                               const double *normals, double *out, double tol, int normal_mode)
{
    orc_line *geo = malloc(sizeof(*geo) * (n_lines ? n_lines : 1));
    for (unsigned il = 0; il < n_lines; ++il)
        geo[il] = orc_line_geom(lines, pos, il);
#pragma omp parallel
    {
        double *krow = malloc(sizeof(double) * 3 * (n_lines ? n_lines : 1));
        long icp;
#pragma omp for schedule(static)
        for (icp = 0; icp < (long)n_cpts; ++icp)
        {
            for (unsigned il = 0; il < n_lines; ++il)
                orc_pair(geo + il, cpts + 3 * icp, tol, krow + 3 * (size_t)il);
            for (unsigned s = 0; s < n_surf; ++s)
            {
                double r[3];
                orc_reduce(s, off, ent, krow, r);
                if (normal_mode)
                {
                    const double *n = normals + 3 * icp;
                    out[(size_t)icp * n_surf + s] = r[0] * n[0] + r[1] * n[1] + r[2] * n[2];
                }
                else
                {
                    double *o = out + ((size_t)icp * n_surf + s) * 3;
                    o[0] = r[0], o[1] = r[1], o[2] = r[2];
                }
            }
        }
        free(krow);
    }
    free(geo);
}

void orc_induction_matrix3(unsigned n_lines, const uint32_t *lines, unsigned n_surf, const uint32_t *off,
                           const uint32_t *ent, const double *pos, unsigned n_cpts, const double *cpts,
                           const double *normals, double *out, double tol)
{
    orc_surface_common(n_lines, lines, n_surf, off, ent, pos, n_cpts, cpts, normals, out, tol, 1);
}

void orc_induction_matrix(unsigned n_lines, const uint32_t *lines, unsigned n_surf, const uint32_t *off,
                          const uint32_t *ent, const double *pos, unsigned n_cpts, const double *cpts, double *out,
                          double tol)
{
    orc_surface_common(n_lines, lines, n_surf, off, ent, pos, n_cpts, cpts, NULL, out, tol, 0);
}

/* Point-velocity sum the reference forms as np.vecdot(ind, gamma[None,:,None], axis=1)
 * (python/pyvl/wake_models.py:125-126): here the surface sum is taken in surface order,
 * sum_s gamma_s * K_s, from the same surface induction rows. */
void orc_point_velocity(unsigned n_lines, const uint32_t *lines, unsigned n_surf, const uint32_t *off,
                        const uint32_t *ent, const double *pos, const double *circ, unsigned n_pts,
                        const double *pts, double *out, double tol)
{
    orc_line *geo = malloc(sizeof(*geo) * (n_lines ? n_lines : 1));
    for (unsigned il = 0; il < n_lines; ++il)
        geo[il] = orc_line_geom(lines, pos, il);
#pragma omp parallel
    {
        double *krow = malloc(sizeof(double) * 3 * (n_lines ? n_lines : 1));
        long ip;
#pragma omp for schedule(static)
        for (ip = 0; ip < (long)n_pts; ++ip)
        {
            for (unsigned il = 0; il < n_lines; ++il)
                orc_pair(geo + il, pts + 3 * ip, tol, krow + 3 * (size_t)il);
            double v[3] = {0.0, 0.0, 0.0};
            for (unsigned s = 0; s < n_surf; ++s)
            {
                double r[3];
                orc_reduce(s, off, ent, krow, r);
                v[0] = v[0] + r[0] * circ[s];
                v[1] = v[1] + r[1] * circ[s];
                v[2] = v[2] + r[2] * circ[s];
            }
            out[3 * ip + 0] = v[0];
            out[3 * ip + 1] = v[1];
            out[3 * ip + 2] = v[2];
        }
        free(krow);
    }
    free(geo);
}

/* surface_center (src/core/mesh.c:24-46) and surface_normal (src/core/mesh.c:48-68, with
 * line_direction :14-22 and real3_unit src/core/common.h:122-126). */
void orc_surface_center(const uint32_t *lines, unsigned n_surf, const uint32_t *off, const uint32_t *ent,
                        const double *pos, double *out)
{
    for (unsigned s = 0; s < n_surf; ++s)
    {
        double c[3] = {0.0, 0.0, 0.0};
        for (uint32_t k = off[s]; k < off[s + 1]; ++k)
        {
            const uint32_t e = ent[k], l = e & ORC_MASK;
            const uint32_t p = (e & ORC_REV) ? (lines[2 * l + 1] & ORC_MASK) : (lines[2 * l] & ORC_MASK);
            c[0] = c[0] + pos[3 * p + 0];
            c[1] = c[1] + pos[3 * p + 1];
            c[2] = c[2] + pos[3 * p + 2];
        }
        const double div = 1.0 / (double)(off[s + 1] - off[s]);
        out[3 * s + 0] = c[0] * div;
        out[3 * s + 1] = c[1] * div;
        out[3 * s + 2] = c[2] * div;
    }
}

static void orc_dir(const uint32_t *lines, const double *pos, uint32_t e, double *d)
{
    const uint32_t l = e & ORC_MASK;
    const uint32_t a = lines[2 * l] & ORC_MASK, b = lines[2 * l + 1] & ORC_MASK;
    const uint32_t u = (e & ORC_REV) ? b : a, v = (e & ORC_REV) ? a : b;
    d[0] = pos[3 * u + 0] - pos[3 * v + 0];
    d[1] = pos[3 * u + 1] - pos[3 * v + 1];
    d[2] = pos[3 * u + 2] - pos[3 * v + 2];
}

void orc_surface_normal(const uint32_t *lines, unsigned n_surf, const uint32_t *off, const uint32_t *ent,
                        const double *pos, double *out)
{
    for (unsigned s = 0; s < n_surf; ++s)
    {
        double d1[3], d0[3], o[3] = {0.0, 0.0, 0.0};
        orc_dir(lines, pos, ent[off[s]], d1);
        memcpy(d0, d1, sizeof(d0));
        for (uint32_t k = off[s] + 1; k < off[s + 1]; ++k)
        {
            double d2[3];
            orc_dir(lines, pos, ent[k], d2);
            o[0] = o[0] + (d2[1] * d1[2] - d2[2] * d1[1]);
            o[1] = o[1] + (d2[2] * d1[0] - d2[0] * d1[2]);
            o[2] = o[2] + (d2[0] * d1[1] - d2[1] * d1[0]);
            memcpy(d1, d2, sizeof(d1));
        }
        o[0] = o[0] + (d0[1] * d1[2] - d0[2] * d1[1]);
        o[1] = o[1] + (d0[2] * d1[0] - d0[0] * d1[2]);
        o[2] = o[2] + (d0[0] * d1[1] - d0[1] * d1[0]);
        const double inv = 1.0 / sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        out[3 * s + 0] = o[0] * inv;
        out[3 * s + 1] = o[1] * inv;
        out[3 * s + 2] = o[2] * inv;
    }
}

/* mesh_from_elements (src/core/mesh.c:175-247): first-match linear search over the lines
 * made so far.  lines_out holds 2*sum(counts), entries sum(counts). */
int orc_mesh_from_elements(unsigned n_el, const unsigned *counts, const unsigned *flat, unsigned *n_lines_out,
                           uint32_t *lines_out, uint32_t *off_out, uint32_t *ent_out)
{
    unsigned n_lines = 0, off = 0;
    off_out[0] = 0;
    for (unsigned el = 0; el < n_el; ++el)
    {
        const unsigned n = counts[el];
        unsigned left = n ? flat[off + n - 1] : 0;
        for (unsigned k = 0; k < n; ++k)
        {
            const unsigned right = flat[off + k];
            unsigned il, orient = 0;
            for (il = 0; il < n_lines; ++il)
            {
                if (lines_out[2 * il] == left && lines_out[2 * il + 1] == right)
                {
                    orient = 0;
                    break;
                }
                if (lines_out[2 * il + 1] == left && lines_out[2 * il] == right)
                {
                    orient = 1;
                    break;
                }
            }
            ent_out[off + k] = il | (orient << 31);
            if (il == n_lines)
            {
                lines_out[2 * n_lines] = left;
                lines_out[2 * n_lines + 1] = right;
                n_lines += 1;
            }
            left = right;
        }
        off += n;
        off_out[el + 1] = off;
    }
    *n_lines_out = n_lines;
    return 0;
}

/* mesh_dual_from_primal (src/core/mesh.c:70-173). dual_lines[2*n_lines];
 * dual_off[n_points+1]; dual_ent[2*n_lines]. */
int orc_mesh_dual(unsigned n_points, unsigned n_lines, const uint32_t *lines, unsigned n_surf, const uint32_t *off,
                  const uint32_t *ent, uint32_t *dual_lines, uint32_t *dual_off, uint32_t *dual_ent)
{
    for (unsigned il = 0; il < n_lines; ++il)
    {
        uint32_t ids[2] = {ORC_INVALID, ORC_INVALID};
        unsigned cnt = 0;
        for (unsigned s = 0; s < n_surf && cnt < 2; ++s)
        {
            for (uint32_t k = off[s]; k < off[s + 1]; ++k)
            {
                if ((ent[k] & ORC_MASK) == il)
                {
                    ids[(ent[k] & ORC_REV) ? 0 : 1] = s;
                    cnt += 1;
                    break;
                }
            }
        }
        dual_lines[2 * il] = ids[0];
        dual_lines[2 * il + 1] = ids[1];
    }
    unsigned o = 0;
    dual_off[0] = 0;
    for (unsigned p = 0; p < n_points; ++p)
    {
        for (unsigned il = 0; il < n_lines; ++il)
        {
            if ((lines[2 * il] & ORC_MASK) == p)
                dual_ent[o++] = il;
            if ((lines[2 * il + 1] & ORC_MASK) == p)
                dual_ent[o++] = il | ORC_REV;
        }
        dual_off[p + 1] = o;
    }
    return 0;
}

/* Plain right-looking FP64 LU with partial pivoting (LAPACK dgetf2 semantics: first index
 * of the largest |a| in the column, row swaps applied to the whole row), row-major A. Used
 * only as a cross-check of the library solver on tiny systems; the reference's solver is
 * LAPACK dgetrf through scipy (python/pyvl/solver.py:223-227). */
int orc_lu_factor(unsigned n, double *a, int *piv)
{
    for (unsigned j = 0; j < n; ++j)
    {
        unsigned p = j;
        double best = fabs(a[(size_t)j * n + j]);
        for (unsigned i = j + 1; i < n; ++i)
        {
            const double v = fabs(a[(size_t)i * n + j]);
            if (v > best)
            {
                best = v;
                p = i;
            }
        }
        piv[j] = (int)p;
        if (a[(size_t)p * n + j] == 0.0)
            return (int)j + 1;
        if (p != j)
            for (unsigned c = 0; c < n; ++c)
            {
                const double t = a[(size_t)j * n + c];
                a[(size_t)j * n + c] = a[(size_t)p * n + c];
                a[(size_t)p * n + c] = t;
            }
        const double inv = 1.0 / a[(size_t)j * n + j];
        for (unsigned i = j + 1; i < n; ++i)
        {
            const double l = a[(size_t)i * n + j] * inv;
            a[(size_t)i * n + j] = l;
            for (unsigned c = j + 1; c < n; ++c)
                a[(size_t)i * n + c] = a[(size_t)i * n + c] - l * a[(size_t)j * n + c];
        }
    }
    return 0;
}

/* Mesh.line_forces (src/meshobject.c:1350-1393): for primal line l = (p1, p2) with dual line
 * (s+, s-), the line circulation c = (s+ valid ? circ[s+] : 0) - (s- valid ? circ[s-] : 0)
 * and f_l = (r_p2 - r_p1) x ((v_p1 + v_p2) * (0.5 c)), with the reference's real3_add,
 * real3_mul1 and real3_cross (src/core/common.h:98-136) in their operation order. */
void orc_line_forces(unsigned n_lines, const uint32_t *primal_lines, const uint32_t *dual_lines, const double *circ,
                     const double *pos, const double *vel, double *f)
{
    for (unsigned il = 0; il < n_lines; ++il)
    {
        const uint32_t p1 = primal_lines[2 * il] & ORC_MASK, p2 = primal_lines[2 * il + 1] & ORC_MASK;
        const uint32_t d1 = dual_lines[2 * il] & ORC_MASK, d2 = dual_lines[2 * il + 1] & ORC_MASK;
        double dr[3], av[3];
        for (int c = 0; c < 3; ++c)
            dr[c] = pos[3 * p2 + c] - pos[3 * p1 + c];
        double lc = 0;
        if (d1 != ORC_INVALID)
            lc += circ[d1];
        if (d2 != ORC_INVALID)
            lc -= circ[d2];
        const double k = 0.5 * lc;
        for (int c = 0; c < 3; ++c)
            av[c] = (vel[3 * p1 + c] + vel[3 * p2 + c]) * k;
        f[3 * il + 0] = dr[1] * av[2] - dr[2] * av[1];
        f[3 * il + 1] = dr[2] * av[0] - dr[0] * av[2];
        f[3 * il + 2] = dr[0] * av[1] - dr[1] * av[0];
    }
}
