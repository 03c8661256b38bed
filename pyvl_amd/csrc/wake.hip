// Unsteady line wake of python/pyvl/wake_models.py (WakeModelLineExplicitUnsteady), resident
// on the GPU.
//
// Reference state: wake_positions (n_shed, 2, R+1, 3) and circulation (n_shed, R); the ring
// mesh of _create_wake_mesh (:35-52) gives every strip its own nodes, so neighbouring strips
// carry coincident nodes and coincident opposite filaments.  Here the state is stored once
// per distinct node, row-major:
//   upos[r * n_te + c]   node of wake row r at trailing-edge column c (c indexes the distinct
//                        bordering nodes, in first-appearance order over (strip, side));
//   circ[r * n_shed + s] circulation of ring (s, r).
// Coincident reference nodes have identical positions at every step (same start point, same
// velocity), so the reference layout is recovered exactly by a gather (pvl_wake_get_state).
// Rings are numbered row-major and lines in first-appearance order over that ring list, so
// the lines of the first r ring rows are a prefix: kernels touch only the filled rows
// (unfilled reference rows are collapsed rings with zero circulation and contribute 0).
//
// update (wake_models.py:83-157), per step:
//   old rows 0..cr-1 (cr = min(step_count, R)) are convected by explicit Euler into rows
//   1..cr with v = V_inf + (bound induction) [+ wake self-induction in free-wake mode];
//   row 0 is pinned to the trailing-edge nodes; rows > cr copy row cr; circulation rows shift
//   by one and row 0 = Gamma[adj1] - Gamma[adj0] (INVALID_ID sides skipped).
// apply_corrections (:177-223): ring row 0 enters the matrix (A[:, adj1] -= w, A[:, adj0]
//   += w per shed line, in shed-line order), rings of rows >= 1 enter the right-hand side
//   b += n . sum Gamma K.
#include "pvl_internal.hpp"

#include <algorithm>
#include <unordered_map>
#include <vector>

using namespace pvl;

struct pvl_wake
{
    unsigned n_shed = 0, R = 0, n_te = 0;
    double tol = 0, current_time = 0;
    unsigned step_count = 0;
    std::vector<uint32_t> col;         // [n_shed*2] column of (strip, side)
    std::vector<uint32_t> te_nodes;    // [n_te] bound node per column
    std::vector<uint32_t> nl_prefix;   // [R+1] lines used by ring rows < r
    std::vector<uint32_t> adj;         // [n_shed*2] adjacent surfaces (host copy)
    uint32_t *d_te_nodes = nullptr;    // [n_te]
    uint32_t *d_adj = nullptr;         // [n_shed*2]
    double *d_upos = nullptr;          // [(R+1)*n_te*3]
    double *d_circ[2] = {nullptr, nullptr}; // [R*n_shed]
    int cur = 0;
    double *d_old = nullptr;           // [R*n_te*3]
    double *d_vel = nullptr;           // [R*n_te*3]
    double *d_circ_x = nullptr;        // [R*n_shed] explicit-part circulation (row 0 zeroed)
    pvl_mesh *full = nullptr;          // all rings, row-major
    pvl_mesh *row0 = nullptr;          // ring row 0 only
    // Self-induction partial sums started early on the side stream (pvl_wake_begin_update),
    // consumed by the next pvl_wake_update.
    void *d_sgeom = nullptr;           // [full->n_lines] LineGeom
    double *d_sgamma = nullptr;        // [full->n_lines]
    double *d_spartial = nullptr;      // [3 * n_pts * chunks]
    size_t spartial_bytes = 0;
    hipEvent_t ev_ready = nullptr, ev_done = nullptr;
    // The self-induction's segments as chains over the distinct nodes (chain_velocity_kernel):
    // streamwise chains per trailing-edge column, spanwise chains per row along runs of
    // adjacent strips, cut into pieces ordered by the ring row that activates them.
    bool chains = false;
    unsigned n_seg = 0;
    uint32_t *d_seg_ab = nullptr;   // [2 n_seg] node pairs, chain direction
    uint32_t *d_seg_line = nullptr; // [n_seg] line of `full` | kReversed when against it
    ChainPiece *d_pieces = nullptr;
    std::vector<uint32_t> pieces_at; // [R+1] pieces active with r ring rows filled
    SegRec *d_segrec = nullptr;      // [n_seg] per-step records
    bool pending = false;
    int pending_prec = 0;
    unsigned pending_na = 0, pending_npts = 0, pending_chunks = 0;
};

namespace
{

__global__ void pin_row0_kernel(double *__restrict__ upos, unsigned n_te, const double *__restrict__ bound_pos,
                                const uint32_t *__restrict__ te_nodes)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_te)
        return;
    const uint32_t p = te_nodes[i];
    upos[3 * i + 0] = bound_pos[3 * p + 0];
    upos[3 * i + 1] = bound_pos[3 * p + 1];
    upos[3 * i + 2] = bound_pos[3 * p + 2];
}

// upos[r+1] = old[r] + (V_inf(old[r]) + vel[r]) * dt for rows r < cr
// (python/pyvl/wake_models.py:121-128: the free stream at the node's position before the step)
__global__ void advect_kernel(double *__restrict__ upos, const double *__restrict__ old,
                              const double *__restrict__ vel, unsigned n, unsigned n_te, FlowField flow, double dt)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double *o = old + 3 * i, *v = vel + 3 * i;
    double *d = upos + 3 * ((size_t)n_te + i);
    double vx, vy, vz;
    flow_velocity(flow, o[0], o[1], o[2], vx, vy, vz);
    d[0] = o[0] + (vx + v[0]) * dt;
    d[1] = o[1] + (vy + v[1]) * dt;
    d[2] = o[2] + (vz + v[2]) * dt;
}

// rows cr+1..R copy row cr (python/pyvl/wake_models.py:134-137)
__global__ void fill_rows_kernel(double *__restrict__ upos, unsigned n_te, unsigned cr, unsigned R)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned n = (R - cr) * n_te;
    if (i >= n)
        return;
    const unsigned r = cr + 1 + i / n_te, c = i % n_te;
    for (int k = 0; k < 3; ++k)
        upos[3 * ((size_t)r * n_te + c) + k] = upos[3 * ((size_t)cr * n_te + c) + k];
}

// circulation shift and shedding (python/pyvl/wake_models.py:139-151)
__global__ void shed_kernel(const double *__restrict__ old, double *__restrict__ neu, unsigned n_shed, unsigned R,
                            const uint32_t *__restrict__ adj, const double *__restrict__ gamma)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_shed * R)
        return;
    if (i >= n_shed)
    {
        neu[i] = old[i - n_shed];
        return;
    }
    const uint32_t a0 = adj[2 * i] & kInvalidId, a1 = adj[2 * i + 1] & kInvalidId;
    double g = 0.0;
    if (a1 != kInvalidId)
        g = g + gamma[a1];
    if (a0 != kInvalidId)
        g = g - gamma[a0];
    neu[i] = g;
}

// A[i][adj1] -= w[i][s], A[i][adj0] += w[i][s] in shed-line order (wake_models.py:209-216)
__global__ void implicit_kernel(double *__restrict__ A, unsigned n_cpts, unsigned n_surf,
                                const double *__restrict__ w, unsigned n_shed, const uint32_t *__restrict__ adj)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_cpts)
        return;
    double *row = A + (size_t)i * n_surf;
    const double *wr = w + (size_t)i * n_shed;
    for (unsigned s = 0; s < n_shed; ++s)
    {
        const uint32_t a0 = adj[2 * s] & kInvalidId, a1 = adj[2 * s + 1] & kInvalidId;
        const double c = wr[s];
        if (a1 != kInvalidId)
            row[a1] = row[a1] - c;
        if (a0 != kInvalidId)
            row[a0] = row[a0] + c;
    }
}

// b[i] += n_i . v_i  (explicit part, wake_models.py:218-223)
__global__ void explicit_kernel(double *__restrict__ b, const double *__restrict__ normals,
                                const double *__restrict__ v, unsigned n_cpts)
{
    main_stream_priority();
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_cpts)
        return;
    const double *n = normals + 3 * i, *vv = v + 3 * i;
    b[i] = b[i] + (vv[0] * n[0] + vv[1] * n[1] + vv[2] * n[2]);
}

int build_topology(pvl_wake *w, std::vector<uint32_t> &lines, std::vector<uint32_t> &off,
                   std::vector<uint32_t> &ent, unsigned rows, std::vector<uint32_t> *prefix)
{
    // Ring (s, r) in row-major order, nodes [u(r,s,0), u(r,s,1), u(r+1,s,1), u(r+1,s,0)],
    // lines deduplicated exactly as pvl_mesh_from_elements does.
    const unsigned n_el = rows * w->n_shed;
    std::vector<unsigned> counts(n_el, 4), flat(4 * (size_t)n_el);
    for (unsigned r = 0; r < rows; ++r)
        for (unsigned s = 0; s < w->n_shed; ++s)
        {
            const size_t e = (size_t)r * w->n_shed + s;
            const uint32_t c0 = w->col[2 * s], c1 = w->col[2 * s + 1];
            flat[4 * e + 0] = r * w->n_te + c0;
            flat[4 * e + 1] = r * w->n_te + c1;
            flat[4 * e + 2] = (r + 1) * w->n_te + c1;
            flat[4 * e + 3] = (r + 1) * w->n_te + c0;
        }
    lines.assign(2 * flat.size(), 0);
    off.assign(n_el + 1, 0);
    ent.assign(flat.size(), 0);
    unsigned nl = 0;
    if (prefix)
    {
        // same construction row by row to record the line-count prefix
        std::unordered_map<uint64_t, uint32_t> seen;
        seen.reserve(flat.size());
        prefix->assign(rows + 1, 0);
        size_t k = 0;
        for (unsigned r = 0; r < rows; ++r)
        {
            for (unsigned s = 0; s < w->n_shed; ++s)
            {
                const size_t e = (size_t)r * w->n_shed + s;
                uint32_t left = flat[4 * e + 3];
                for (int q = 0; q < 4; ++q)
                {
                    const uint32_t right = flat[4 * e + q];
                    const uint32_t lo = left < right ? left : right, hi = left < right ? right : left;
                    const uint64_t key = (uint64_t(hi) << 32) | lo;
                    auto it = seen.find(key);
                    uint32_t id, orient = 0;
                    if (it == seen.end())
                    {
                        id = nl++;
                        seen.emplace(key, id);
                        lines[2 * id] = left;
                        lines[2 * id + 1] = right;
                    }
                    else
                    {
                        id = it->second;
                        orient = (lines[2 * id] == left && lines[2 * id + 1] == right) ? 0u : 1u;
                    }
                    ent[k++] = id | (orient << 31);
                    left = right;
                }
                off[e + 1] = (uint32_t)k;
            }
            (*prefix)[r + 1] = nl;
        }
    }
    else
    {
        if (int st = pvl_mesh_from_elements(n_el, counts.data(), flat.data(), &nl, lines.data(), off.data(),
                                            ent.data()))
            return st;
    }
    lines.resize(2 * (size_t)nl);
    return PVL_OK;
}

// The chains of the wake's lines (see pvl_wake::chains).  Every line of `full` must be one
// segment of one chain; otherwise (unusual strip topologies) the generic line sum is used.
int build_chains(pvl_wake *w, const std::vector<uint32_t> &lines)
{
    const unsigned R = w->R, n_te = w->n_te, n_shed = w->n_shed;
    const size_t n_lines = lines.size() / 2;
    std::unordered_map<uint64_t, uint32_t> id;
    id.reserve(n_lines);
    for (size_t l = 0; l < n_lines; ++l)
    {
        const uint32_t a = lines[2 * l], b = lines[2 * l + 1];
        id.emplace((uint64_t(a > b ? a : b) << 32) | (a > b ? b : a), (uint32_t)l);
    }
    std::vector<uint32_t> ab, sl;
    std::vector<std::pair<uint32_t, ChainPiece>> keyed;
    std::vector<char> used(n_lines, 0);
    bool ok = true;
    // nodes n_0..n_k of one chain; ring row of segment j: streamwise j, spanwise (row - 1, row)
    auto add_chain = [&](const std::vector<uint32_t> &nodes, bool streamwise, unsigned row) {
        const unsigned k = (unsigned)nodes.size() - 1;
        for (unsigned q = 0; q < k; q += kChainPiece)
        {
            ChainPiece pc;
            pc.start_node = nodes[q];
            pc.first_seg = (uint32_t)sl.size();
            pc.n_seg = (k - q) < kChainPiece ? (k - q) : kChainPiece;
            pc.row1 = streamwise ? q + 1 : 0;
            const uint32_t key = streamwise ? q + 1 : (row == 0 ? 1 : row);
            keyed.push_back({key, pc});
            for (unsigned j = q; j < q + pc.n_seg; ++j)
            {
                const uint32_t a = nodes[j], b = nodes[j + 1];
                auto it = id.find((uint64_t(a > b ? a : b) << 32) | (a > b ? b : a));
                if (it == id.end() || used[it->second])
                {
                    ok = false;
                    return;
                }
                used[it->second] = 1;
                const uint32_t l = it->second;
                ab.push_back(a), ab.push_back(b);
                sl.push_back(lines[2 * l] == a ? l : (l | kReversed));
            }
        }
    };
    // runs of adjacent strips: strip t follows strip s when its side-0 column is s's side-1 column
    std::vector<std::vector<uint32_t>> starting(n_te);
    std::vector<char> has_pred(n_shed, 0), seen(n_shed, 0);
    for (unsigned s2 = 0; s2 < n_shed; ++s2)
        starting[w->col[2 * s2]].push_back(s2);
    for (unsigned s2 = 0; s2 < n_shed; ++s2)
        for (uint32_t t : starting[w->col[2 * s2 + 1]])
            if (t != s2)
                has_pred[t] = 1;
    std::vector<std::vector<uint32_t>> runs;
    for (int pass = 0; pass < 2; ++pass)
        for (unsigned s0 = 0; s0 < n_shed; ++s0)
        {
            if (seen[s0] || (pass == 0 && has_pred[s0]))
                continue;
            std::vector<uint32_t> run;
            for (int cur = (int)s0; cur >= 0;)
            {
                seen[cur] = 1;
                run.push_back((uint32_t)cur);
                int nxt = -1;
                for (uint32_t t : starting[w->col[2 * cur + 1]])
                    if (!seen[t])
                    {
                        nxt = (int)t;
                        break;
                    }
                cur = nxt;
            }
            runs.push_back(run);
        }
    for (unsigned r = 0; r <= R && ok; ++r)
        for (const auto &run : runs)
        {
            std::vector<uint32_t> nodes{r * n_te + w->col[2 * run[0]]};
            for (uint32_t s2 : run)
                nodes.push_back(r * n_te + w->col[2 * s2 + 1]);
            add_chain(nodes, false, r);
            if (!ok)
                break;
        }
    for (unsigned c = 0; c < n_te && ok; ++c)
    {
        std::vector<uint32_t> nodes(R + 1);
        for (unsigned r = 0; r <= R; ++r)
            nodes[r] = r * n_te + c;
        add_chain(nodes, true, 0);
    }
    if (!ok || sl.size() != n_lines)
        return PVL_OK; // generic path
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const std::pair<uint32_t, ChainPiece> &x, const std::pair<uint32_t, ChainPiece> &y) {
                         return x.first < y.first;
                     });
    std::vector<ChainPiece> pieces(keyed.size());
    w->pieces_at.assign(R + 1, 0);
    for (size_t i = 0; i < keyed.size(); ++i)
        pieces[i] = keyed[i].second;
    for (unsigned rr = 0; rr <= R; ++rr)
    {
        uint32_t cnt = 0;
        while (cnt < keyed.size() && keyed[cnt].first <= rr)
            ++cnt;
        w->pieces_at[rr] = cnt;
    }
    w->n_seg = (unsigned)sl.size();
    PVL_HIP(hipMalloc((void **)&w->d_seg_ab, sizeof(uint32_t) * ab.size()));
    PVL_HIP(hipMalloc((void **)&w->d_seg_line, sizeof(uint32_t) * sl.size()));
    PVL_HIP(hipMalloc((void **)&w->d_pieces, sizeof(ChainPiece) * pieces.size()));
    PVL_HIP(hipMalloc((void **)&w->d_segrec, sizeof(SegRec) * sl.size()));
    PVL_HIP(hipMemcpy(w->d_seg_ab, ab.data(), sizeof(uint32_t) * ab.size(), hipMemcpyHostToDevice));
    PVL_HIP(hipMemcpy(w->d_seg_line, sl.data(), sizeof(uint32_t) * sl.size(), hipMemcpyHostToDevice));
    PVL_HIP(hipMemcpy(w->d_pieces, pieces.data(), sizeof(ChainPiece) * pieces.size(), hipMemcpyHostToDevice));
    w->chains = true;
    return PVL_OK;
}

} // namespace

extern "C" {

int pvl_wake_destroy(pvl_wake *w);

int pvl_wake_create(pvl_wake **out, unsigned n_shed, unsigned line_rows, const uint32_t *bordering_nodes,
                    const uint32_t *adjacent_surfaces, double vortex_tol, double time)
{
    PVL_API_LOCK;
    *out = nullptr;
    if (int st = ensure_init())
        return st;
    if (n_shed == 0)
    {
        set_error("The geometry has no edges which would shed any wake.");
        return PVL_ERR_ARG;
    }
    if (line_rows == 0)
    {
        set_error("line_rows must be positive");
        return PVL_ERR_ARG;
    }
    pvl_wake *w = new pvl_wake;
    w->n_shed = n_shed;
    w->R = line_rows;
    w->tol = vortex_tol;
    w->current_time = time;
    w->col.resize(2 * (size_t)n_shed);
    std::unordered_map<uint32_t, uint32_t> colmap;
    for (unsigned i = 0; i < 2 * n_shed; ++i)
    {
        const uint32_t node = bordering_nodes[i];
        auto it = colmap.find(node);
        if (it == colmap.end())
        {
            it = colmap.emplace(node, (uint32_t)w->te_nodes.size()).first;
            w->te_nodes.push_back(node);
        }
        w->col[i] = it->second;
    }
    w->n_te = (unsigned)w->te_nodes.size();
    w->adj.assign(adjacent_surfaces, adjacent_surfaces + 2 * (size_t)n_shed);
    const unsigned R = line_rows, n_te = w->n_te;
    std::vector<uint32_t> lines, off, ent;
    int st = build_topology(w, lines, off, ent, R, &w->nl_prefix);
    if (!st)
        st = pvl_mesh_create(&w->full, (R + 1) * n_te, (unsigned)(lines.size() / 2), lines.data(), R * n_shed,
                             off.data(), ent.data());
    if (!st)
        st = build_chains(w, lines);
    if (!st)
    {
        std::vector<uint32_t> l0, o0, e0;
        st = build_topology(w, l0, o0, e0, 1, nullptr);
        if (!st)
            st = pvl_mesh_create(&w->row0, (R + 1) * n_te, (unsigned)(l0.size() / 2), l0.data(), n_shed, o0.data(),
                                 e0.data());
    }
    auto dalloc = [&](void **p, size_t bytes) {
        if (st)
            return;
        st = check_hip(hipMalloc(p, bytes ? bytes : 8), "hipMalloc(wake)");
    };
    dalloc((void **)&w->d_te_nodes, sizeof(uint32_t) * n_te);
    dalloc((void **)&w->d_adj, sizeof(uint32_t) * 2 * n_shed);
    dalloc((void **)&w->d_upos, sizeof(double) * 3 * (size_t)(R + 1) * n_te);
    dalloc((void **)&w->d_circ[0], sizeof(double) * (size_t)R * n_shed);
    dalloc((void **)&w->d_circ[1], sizeof(double) * (size_t)R * n_shed);
    dalloc((void **)&w->d_old, sizeof(double) * 3 * (size_t)R * n_te);
    dalloc((void **)&w->d_vel, sizeof(double) * 3 * (size_t)R * n_te);
    dalloc((void **)&w->d_circ_x, sizeof(double) * (size_t)R * n_shed);
    if (!st)
        st = check_hip(hipMemcpy(w->d_te_nodes, w->te_nodes.data(), sizeof(uint32_t) * n_te, hipMemcpyHostToDevice),
                       "hipMemcpy");
    if (!st)
        st = check_hip(hipMemcpy(w->d_adj, adjacent_surfaces, sizeof(uint32_t) * 2 * n_shed, hipMemcpyHostToDevice),
                       "hipMemcpy");
    if (!st)
        st = check_hip(hipMemset(w->d_upos, 0, sizeof(double) * 3 * (size_t)(R + 1) * n_te), "hipMemset");
    if (!st)
        st = check_hip(hipMemset(w->d_circ[0], 0, sizeof(double) * (size_t)R * n_shed), "hipMemset");
    if (st)
    {
        pvl_wake_destroy(w);
        return st;
    }
    *out = w;
    return PVL_OK;
}

int pvl_wake_destroy(pvl_wake *w)
{
    PVL_API_LOCK;
    if (!w)
        return PVL_OK;
    (void)hipStreamSynchronize(stream());
    if (w->ev_done)
    {
        (void)hipEventSynchronize(w->ev_done);
        (void)hipEventDestroy(w->ev_done);
    }
    if (w->ev_ready)
        (void)hipEventDestroy(w->ev_ready);
    (void)hipFree(w->d_sgeom);
    (void)hipFree(w->d_sgamma);
    (void)hipFree(w->d_spartial);
    (void)hipFree(w->d_seg_ab);
    (void)hipFree(w->d_seg_line);
    (void)hipFree(w->d_pieces);
    (void)hipFree(w->d_segrec);
    pvl_mesh_destroy(w->full);
    pvl_mesh_destroy(w->row0);
    (void)hipFree(w->d_te_nodes);
    (void)hipFree(w->d_adj);
    (void)hipFree(w->d_upos);
    (void)hipFree(w->d_circ[0]);
    (void)hipFree(w->d_circ[1]);
    (void)hipFree(w->d_old);
    (void)hipFree(w->d_vel);
    (void)hipFree(w->d_circ_x);
    delete w;
    return PVL_OK;
}

int pvl_wake_info(const pvl_wake *w, unsigned *n_te, unsigned *n_lines, unsigned *step_count, double *current_time)
{
    PVL_API_LOCK;
    if (n_te)
        *n_te = w->n_te;
    if (n_lines)
        *n_lines = w->full->n_lines;
    if (step_count)
        *step_count = w->step_count;
    if (current_time)
        *current_time = w->current_time;
    return PVL_OK;
}

// Ring rows that may carry circulation: min(step_count - 1, R).
static unsigned active_rows(const pvl_wake *w)
{
    if (w->step_count <= 1)
        return 0;
    return w->step_count - 1 < w->R ? w->step_count - 1 : w->R;
}

// Lines carrying circulation: those of ring rows < active_rows.
static unsigned active_lines(const pvl_wake *w) { return w->nl_prefix[active_rows(w)]; }

static unsigned self_chunks(const pvl_wake *w, unsigned n_pts)
{
    if (w->chains)
        return chain_velocity_chunks(n_pts, w->pieces_at[active_rows(w)]);
    return point_velocity_chunks(n_pts, active_lines(w), "wake_self_velocity");
}

// Buffers of the self-induction: line circulations, segment records (chains) or line
// geometry, and the per-chunk partial sums (grown on demand; a started sum must have been
// consumed or dropped before they are regrown).
static int self_buffers(pvl_wake *w, unsigned n_pts)
{
    if (!w->chains && !w->d_sgeom)
        PVL_HIP(hipMalloc(&w->d_sgeom, line_geom_bytes(w->full->n_lines)));
    if (!w->d_sgamma)
        PVL_HIP(hipMalloc((void **)&w->d_sgamma, sizeof(double) * (w->full->n_lines ? w->full->n_lines : 1)));
    const size_t pbytes = sizeof(double) * 3 * (size_t)n_pts * self_chunks(w, n_pts);
    if (pbytes > w->spartial_bytes)
    {
        // no queued kernel on either library stream may still use the old buffer
        PVL_HIP(hipStreamSynchronize(stream()));
        if (hipStream_t s2 = side_stream())
            PVL_HIP(hipStreamSynchronize(s2));
        (void)hipFree(w->d_spartial);
        w->d_spartial = nullptr, w->spartial_bytes = 0;
        PVL_HIP(hipMalloc((void **)&w->d_spartial, pbytes));
        w->spartial_bytes = pbytes;
    }
    return PVL_OK;
}

// Partial sums of the wake's own induction at the n_pts points pts (the first n_pts wake
// nodes), from the state at this moment, on stream s with at most max_groups workgroups.
static int self_partials(pvl_wake *w, unsigned n_pts, const double *pts, int precision, hipStream_t s,
                         unsigned max_groups)
{
    const unsigned ar = active_rows(w), na = w->nl_prefix[ar];
    if (!w->chains)
        return point_velocity_partials(w->full, w->tol, w->d_upos, w->d_circ[w->cur], n_pts, pts, precision, na,
                                       "wake_self_velocity", s, w->d_sgeom, w->d_sgamma, w->d_spartial, max_groups);
    if (int st = line_circulations(w->full, na, w->d_circ[w->cur], w->d_sgamma, s))
        return st;
    if (int st = chain_segments(w->n_seg, w->d_seg_ab, w->d_seg_line, w->d_upos, w->d_sgamma, na, w->tol, w->d_segrec,
                                s))
        return st;
    return chain_velocity_partials(n_pts, pts, w->d_upos, w->d_pieces, w->pieces_at[ar], ar, w->d_segrec, precision,
                                   w->tol, (double)n_pts * (double)na, s, w->d_spartial, max_groups);
}

// Orders the library stream after a started self-induction (its inputs may be overwritten
// from here on) and drops it.
static int drop_pending(pvl_wake *w)
{
    if (!w->pending)
        return PVL_OK;
    w->pending = false;
    PVL_HIP(hipStreamWaitEvent(stream(), w->ev_done, 0));
    return PVL_OK;
}

int pvl_wake_begin_update(pvl_wake *w, int mode, int precision)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (mode != 0 && mode != 1)
    {
        set_error("wake mode must be 0 (reference) or 1 (free wake)");
        return PVL_ERR_ARG;
    }
    if (precision != PVL_FP64 && precision != PVL_FP32 && precision != PVL_FP64_REF)
    {
        set_error("unknown precision %d", precision);
        return PVL_ERR_ARG;
    }
    const unsigned na = active_lines(w);
    if (mode != 1 || na == 0 || w->pending)
        return PVL_OK;
    const unsigned cr = w->step_count < w->R ? w->step_count : w->R;
    const unsigned n_old = cr * w->n_te;
    hipStream_t s = side_stream();
    if (!s)
        return PVL_ERR_HIP;
    if (!w->ev_ready)
        PVL_HIP(hipEventCreateWithFlags(&w->ev_ready, hipEventDisableTiming));
    if (!w->ev_done)
        PVL_HIP(hipEventCreateWithFlags(&w->ev_done, hipEventDisableTiming));
    // no started sum is pending here: the buffers are free to regrow
    if (int st = self_buffers(w, n_old))
        return st;
    const unsigned nc = self_chunks(w, n_old);
    // the wake state as the library stream leaves it now; the points are the first n_old
    // wake nodes (pvl_wake_update evaluates at a copy of the same values)
    PVL_HIP(hipEventRecord(w->ev_ready, stream()));
    PVL_HIP(hipStreamWaitEvent(s, w->ev_ready, 0));
    if (int st = self_partials(w, n_old, w->d_upos, precision, s, (unsigned)side_groups(5)))
        return st;
    PVL_HIP(hipEventRecord(w->ev_done, s));
    w->pending = true;
    w->pending_prec = precision, w->pending_na = na, w->pending_npts = n_old, w->pending_chunks = nc;
    return PVL_OK;
}

int pvl_wake_update(pvl_wake *w, double time, const pvl_mesh *bound, const double *bound_pos_dev,
                    const double *circulation_dev, const double *vinf, int mode, int precision)
{
    const double desc[7] = {0.0, vinf[0], vinf[1], vinf[2], 0.0, 0.0, 0.0};
    return pvl_wake_update_flow(w, time, bound, bound_pos_dev, circulation_dev, desc, mode, precision);
}

int pvl_wake_update_flow(pvl_wake *w, double time, const pvl_mesh *bound, const double *bound_pos_dev,
                         const double *circulation_dev, const double *flow, int mode, int precision)
{
    PVL_API_LOCK;
    FlowField fl;
    if (int st = flow_from_descriptor(flow, &fl))
        return st;
    if (int st = ensure_init())
        return st;
    if (mode != 0 && mode != 1)
    {
        set_error("wake mode must be 0 (reference) or 1 (free wake)");
        return PVL_ERR_ARG;
    }
    const unsigned n_te = w->n_te;
    const double dt = time - w->current_time;
    w->current_time = time;
    if (w->step_count == 0)
    {
        pin_row0_kernel<<<(n_te + 255) / 256, 256, 0, stream()>>>(w->d_upos, n_te, bound_pos_dev, w->d_te_nodes);
        PVL_HIP(hipGetLastError());
        w->step_count = 1;
    }
    const unsigned cr = w->step_count < w->R ? w->step_count : w->R;
    const unsigned n_old = cr * n_te;
    if (int st = device_copy(w->d_old, w->d_upos, sizeof(double) * 3 * n_old))
        return st;
    {
        if (int st = point_velocity_impl(bound, w->tol, bound_pos_dev, circulation_dev, n_old, w->d_old, w->d_vel, 0,
                                         precision, bound->n_lines, "wake_bound_velocity"))
            return st;
    }
    if (mode == 1)
    {
        const unsigned na = active_lines(w);
        if (na && w->pending && w->pending_prec == precision && w->pending_na == na && w->pending_npts == n_old)
        {
            // started by pvl_wake_begin_update: the same partial sums, added in the same order
            w->pending = false;
            PVL_HIP(hipStreamWaitEvent(stream(), w->ev_done, 0));
            if (int st = point_velocity_reduce(n_old, w->pending_chunks, w->d_spartial, w->d_vel, 1, stream()))
                return st;
        }
        else if (na)
        {
            if (int st = drop_pending(w))
                return st;
            if (int st = self_buffers(w, n_old))
                return st;
            if (int st = self_partials(w, n_old, w->d_old, precision, stream(), 0))
                return st;
            if (int st = point_velocity_reduce(n_old, self_chunks(w, n_old), w->d_spartial, w->d_vel, 1, stream()))
                return st;
        }
    }
    if (int st = drop_pending(w))
        return st;
    advect_kernel<<<(n_old + 255) / 256, 256, 0, stream()>>>(w->d_upos, w->d_old, w->d_vel, n_old, n_te, fl, dt);
    PVL_HIP(hipGetLastError());
    pin_row0_kernel<<<(n_te + 255) / 256, 256, 0, stream()>>>(w->d_upos, n_te, bound_pos_dev, w->d_te_nodes);
    PVL_HIP(hipGetLastError());
    if (cr != w->R)
    {
        const unsigned nf = (w->R - cr) * n_te;
        fill_rows_kernel<<<(nf + 255) / 256, 256, 0, stream()>>>(w->d_upos, n_te, cr, w->R);
        PVL_HIP(hipGetLastError());
    }
    const unsigned nc = w->R * w->n_shed;
    shed_kernel<<<(nc + 255) / 256, 256, 0, stream()>>>(w->d_circ[w->cur], w->d_circ[1 - w->cur], w->n_shed, w->R,
                                                        w->d_adj, circulation_dev);
    PVL_HIP(hipGetLastError());
    w->cur = 1 - w->cur;
    w->step_count += 1;
    return PVL_OK;
}

int pvl_wake_apply_corrections(pvl_wake *w, unsigned n_cpts, const double *cpts_dev, const double *normals_dev,
                               double *mat_dev, unsigned n_surf, double *rhs_dev, int precision)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    if (w->step_count == 0 || n_cpts == 0)
        return PVL_OK; // no geometry yet (the reference skips the call at iteration 0)
    for (unsigned i = 0; i < 2 * w->n_shed; ++i)
    {
        const uint32_t a = w->adj[i] & kInvalidId;
        if (a != kInvalidId && a >= n_surf)
        {
            set_error("adjacent surface %u out of range for a %u-column matrix", a, n_surf);
            return PVL_ERR_ARG;
        }
    }
    // implicit part: ring row 0
    double *W0 = (double *)scratch(12, sizeof(double) * (size_t)n_cpts * w->n_shed);
    if (!W0)
        return PVL_ERR_NOMEM;
    {
        if (int st = pvl_induction_matrix3(w->row0, w->tol, w->d_upos, n_cpts, cpts_dev, normals_dev, W0, precision))
            return st;
        implicit_kernel<<<(n_cpts + 255) / 256, 256, 0, stream()>>>(mat_dev, n_cpts, n_surf, W0, w->n_shed,
                                                                     w->d_adj);
        PVL_HIP(hipGetLastError());
    }
    // explicit part: rings of rows >= 1
    const unsigned na = active_lines(w);
    if (na == 0)
        return PVL_OK;
    const size_t nc = (size_t)w->R * w->n_shed;
    if (int st = device_copy(w->d_circ_x, w->d_circ[w->cur], sizeof(double) * nc))
        return st;
    if (int st = device_fill32(w->d_circ_x, 0u, sizeof(double) * w->n_shed))
        return st;
    double *v = (double *)scratch(13, sizeof(double) * 3 * (size_t)n_cpts);
    if (!v)
        return PVL_ERR_NOMEM;
    {
        if (int st = point_velocity_impl(w->full, w->tol, w->d_upos, w->d_circ_x, n_cpts, cpts_dev, v, 0, precision,
                                         na, "wake_explicit"))
            return st;
        explicit_kernel<<<(n_cpts + 255) / 256, 256, 0, stream()>>>(rhs_dev, normals_dev, v, n_cpts);
        PVL_HIP(hipGetLastError());
    }
    return PVL_OK;
}

// WakeModelLineExplicitUnsteady.get_velocity (wake_models.py:159-175).
int pvl_wake_get_velocity(pvl_wake *w, unsigned n_pts, const double *pts_dev, double *out_dev, int accumulate,
                          int precision)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    const unsigned na = active_lines(w);
    if (na == 0)
    {
        if (!accumulate && n_pts)
            if (int st = device_fill32(out_dev, 0u, sizeof(double) * 3 * n_pts))
                return st;
        return PVL_OK;
    }
    return point_velocity_impl(w->full, w->tol, w->d_upos, w->d_circ[w->cur], n_pts, pts_dev, out_dev, accumulate,
                               precision, na);
}

// State in the reference layout: positions (n_shed, 2, R+1, 3), circulation (n_shed, R).
int pvl_wake_get_state(pvl_wake *w, double *positions_host, double *circulation_host)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    const unsigned R = w->R, n_te = w->n_te, n_shed = w->n_shed;
    std::vector<double> up(3 * (size_t)(R + 1) * n_te), ci((size_t)R * n_shed);
    PVL_HIP(hipMemcpyAsync(up.data(), w->d_upos, sizeof(double) * up.size(), hipMemcpyDeviceToHost, stream()));
    PVL_HIP(hipMemcpyAsync(ci.data(), w->d_circ[w->cur], sizeof(double) * ci.size(), hipMemcpyDeviceToHost,
                           stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    if (positions_host)
        for (unsigned s = 0; s < n_shed; ++s)
            for (unsigned sd = 0; sd < 2; ++sd)
                for (unsigned r = 0; r <= R; ++r)
                    for (int k = 0; k < 3; ++k)
                        positions_host[(((size_t)s * 2 + sd) * (R + 1) + r) * 3 + k] =
                            up[((size_t)r * n_te + w->col[2 * s + sd]) * 3 + k];
    if (circulation_host)
        for (unsigned s = 0; s < n_shed; ++s)
            for (unsigned r = 0; r < R; ++r)
                circulation_host[(size_t)s * R + r] = ci[(size_t)r * n_shed + s];
    return PVL_OK;
}

// Restore a saved state (python/pyvl/wake_models.py:253-291).  Positions of coincident
// reference nodes must agree; the first occurrence of each distinct node is used.
int pvl_wake_set_state(pvl_wake *w, const double *positions_host, const double *circulation_host, unsigned step_count,
                       double current_time)
{
    PVL_API_LOCK;
    if (int st = ensure_init())
        return st;
    const unsigned R = w->R, n_te = w->n_te, n_shed = w->n_shed;
    std::vector<double> up(3 * (size_t)(R + 1) * n_te), ci((size_t)R * n_shed);
    std::vector<char> done(n_te, 0);
    for (unsigned s = 0; s < n_shed; ++s)
        for (unsigned sd = 0; sd < 2; ++sd)
        {
            const uint32_t c = w->col[2 * s + sd];
            if (done[c])
                continue;
            done[c] = 1;
            for (unsigned r = 0; r <= R; ++r)
                for (int k = 0; k < 3; ++k)
                    up[((size_t)r * n_te + c) * 3 + k] = positions_host[(((size_t)s * 2 + sd) * (R + 1) + r) * 3 + k];
        }
    for (unsigned s = 0; s < n_shed; ++s)
        for (unsigned r = 0; r < R; ++r)
            ci[(size_t)r * n_shed + s] = circulation_host[(size_t)s * R + r];
    if (int st = drop_pending(w))
        return st;
    PVL_HIP(hipMemcpyAsync(w->d_upos, up.data(), sizeof(double) * up.size(), hipMemcpyHostToDevice, stream()));
    PVL_HIP(hipMemcpyAsync(w->d_circ[w->cur], ci.data(), sizeof(double) * ci.size(), hipMemcpyHostToDevice,
                           stream()));
    PVL_HIP(hipStreamSynchronize(stream()));
    w->step_count = step_count;
    w->current_time = current_time;
    return PVL_OK;
}

} // extern "C"
