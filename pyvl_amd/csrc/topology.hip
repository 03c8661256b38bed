// Mesh topology: construction from element lists, the dual mesh, the closed-form wake
// ring mesh, and the device-resident mesh object.  Host C++ (built by hipcc only so that
// the whole library is one toolchain).
#include "pvl_internal.hpp"

#include <algorithm>
#include <unordered_map>
#include <vector>

using namespace pvl;

namespace
{

inline uint64_t pair_key(uint32_t a, uint32_t b)
{
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    return (uint64_t(hi) << 32) | lo;
}

} // namespace

extern "C" {

// Same result as the reference's O(E*L) first-match search (src/core/mesh.c:175-247):
// each unordered endpoint pair forms at most one line, numbered in order of first
// appearance with p1 = the element's previous point and p2 = the current one; a later
// element walking it in the opposite direction gets the REVERSED bit.  A degenerate
// pair (a, a) always matches with orientation 0, as the reference's first test does.
int pvl_mesh_from_elements(unsigned n_elements, const unsigned *point_counts, const unsigned *flat_points,
                           unsigned *n_lines_out, uint32_t *lines_out, uint32_t *surface_offsets_out,
                           uint32_t *surface_lines_out)
{
    if (n_elements && (!point_counts || !flat_points))
    {
        set_error("mesh_from_elements: null input");
        return PVL_ERR_ARG;
    }
    size_t total = 0;
    for (unsigned i = 0; i < n_elements; ++i)
        total += point_counts[i];
    std::unordered_map<uint64_t, uint32_t> seen;
    seen.reserve(total);
    uint32_t n_lines = 0;
    size_t off = 0;
    surface_offsets_out[0] = 0;
    for (unsigned e = 0; e < n_elements; ++e)
    {
        const unsigned n = point_counts[e];
        if (n == 0)
        {
            surface_offsets_out[e + 1] = (uint32_t)off;
            continue;
        }
        uint32_t left = flat_points[off + n - 1];
        for (unsigned k = 0; k < n; ++k)
        {
            const uint32_t right = flat_points[off + k];
            const uint64_t key = pair_key(left, right);
            auto it = seen.find(key);
            uint32_t id, orient = 0;
            if (it == seen.end())
            {
                id = n_lines++;
                seen.emplace(key, id);
                lines_out[2 * id + 0] = left;
                lines_out[2 * id + 1] = right;
            }
            else
            {
                id = it->second;
                orient = (lines_out[2 * id + 0] == left && lines_out[2 * id + 1] == right) ? 0u : 1u;
            }
            surface_lines_out[off + k] = id | (orient << 31);
            left = right;
        }
        off += n;
        surface_offsets_out[e + 1] = (uint32_t)off;
    }
    *n_lines_out = n_lines;
    return PVL_OK;
}

// Dual mesh with the reference's rules (src/core/mesh.c:70-173): dual line i joins the
// first two surfaces (in surface order) that contain primal line i, the one holding it
// REVERSED in slot p1 and the one holding it forward in slot p2 (only the first
// occurrence inside a surface counts); dual surface k lists, in line order, every primal
// line starting at point k (orientation 0) and then ending at it (orientation 1).
int pvl_mesh_dual(unsigned n_points, unsigned n_lines, const uint32_t *lines, unsigned n_surfaces,
                  const uint32_t *surface_offsets, const uint32_t *surface_lines, uint32_t *dual_lines_out,
                  uint32_t *dual_offsets_out, uint32_t *dual_entries_out)
{
    std::vector<uint8_t> cnt(n_lines, 0);
    std::vector<uint32_t> last_surf(n_lines, 0xFFFFFFFFu);
    for (unsigned i = 0; i < n_lines; ++i)
    {
        dual_lines_out[2 * i + 0] = kInvalidId;
        dual_lines_out[2 * i + 1] = kInvalidId;
    }
    for (unsigned s = 0; s < n_surfaces; ++s)
    {
        for (uint32_t k = surface_offsets[s]; k < surface_offsets[s + 1]; ++k)
        {
            const uint32_t v = surface_lines[k] & kInvalidId;
            const uint32_t o = surface_lines[k] >> 31;
            if (v >= n_lines)
                continue;
            if (last_surf[v] == s)
                continue; // only the first occurrence inside a surface
            last_surf[v] = s;
            if (cnt[v] >= 2)
                continue;
            dual_lines_out[2 * v + (o ? 0 : 1)] = s;
            cnt[v] += 1;
        }
    }
    // counting sort of line ends by point
    std::vector<uint32_t> deg(n_points + 1, 0);
    for (unsigned i = 0; i < n_lines; ++i)
    {
        const uint32_t a = lines[2 * i] & kInvalidId, b = lines[2 * i + 1] & kInvalidId;
        if (a < n_points)
            deg[a] += 1;
        if (b < n_points)
            deg[b] += 1;
    }
    dual_offsets_out[0] = 0;
    for (unsigned p = 0; p < n_points; ++p)
        dual_offsets_out[p + 1] = dual_offsets_out[p] + deg[p];
    std::vector<uint32_t> fill(dual_offsets_out, dual_offsets_out + n_points);
    for (unsigned i = 0; i < n_lines; ++i)
    {
        const uint32_t a = lines[2 * i] & kInvalidId, b = lines[2 * i + 1] & kInvalidId;
        if (a < n_points)
            dual_entries_out[fill[a]++] = i;
        if (b < n_points)
            dual_entries_out[fill[b]++] = i | kReversed;
    }
    return PVL_OK;
}

// Closed form of mesh_from_elements applied to the ring list of
// python/pyvl/wake_models.py:35-52 (ring r of strip s = [r, r+R+1, r+R+2, r+1] + 2(R+1)s):
// ring 0 creates four lines, every later ring three, and reuses the previous ring's
// closing rung reversed.  Line base of strip s is s*(3R+1).
int pvl_wake_topology(unsigned n_strips, unsigned line_rows, unsigned *n_lines_out, uint32_t *lines_out,
                      uint32_t *surface_offsets_out, uint32_t *surface_lines_out)
{
    if (line_rows == 0)
    {
        set_error("wake_topology: line_rows must be positive");
        return PVL_ERR_ARG;
    }
    const uint32_t R = line_rows;
    const uint32_t per_strip = 3 * R + 1;
    for (uint32_t s = 0; s < n_strips; ++s)
    {
        const uint32_t b = 2 * (R + 1) * s; // first node of the strip
        const uint32_t B = per_strip * s;   // first line of the strip
        // ring 0: (e,a), (a,c), (c,d), (d,e)
        uint32_t *L = lines_out + 2 * B;
        L[0] = b + 1, L[1] = b;
        L[2] = b, L[3] = b + R + 1;
        L[4] = b + R + 1, L[5] = b + R + 2;
        L[6] = b + R + 2, L[7] = b + 1;
        for (uint32_t r = 1; r < R; ++r)
        {
            uint32_t *Lr = lines_out + 2 * (B + 3 * r + 1);
            Lr[0] = b + r + 1, Lr[1] = b + r;
            Lr[2] = b + r + R + 1, Lr[3] = b + r + R + 2;
            Lr[4] = b + r + R + 2, Lr[5] = b + r + 1;
        }
        for (uint32_t r = 0; r < R; ++r)
        {
            const uint32_t surf = s * R + r;
            surface_offsets_out[surf] = 4 * surf;
            uint32_t *E = surface_lines_out + 4 * surf;
            if (r == 0)
            {
                E[0] = B, E[1] = B + 1, E[2] = B + 2, E[3] = B + 3;
            }
            else
            {
                E[0] = B + 3 * r + 1;
                E[1] = (B + 3 * r) | kReversed;
                E[2] = B + 3 * r + 2;
                E[3] = B + 3 * r + 3;
            }
        }
    }
    surface_offsets_out[n_strips * R] = 4 * n_strips * R;
    *n_lines_out = n_strips * per_strip;
    return PVL_OK;
}

} // extern "C"

// Assembly program of a mesh (see kAsmSurfaces): per run of consecutive surfaces, entries in
// CSR order; a line used again later in the run is kept in an LDS slot, slots evicted by
// farthest next use (a line whose slot was taken is evaluated again).
void pvl::build_assembly_program(unsigned n_surfaces, const uint32_t *offsets, const uint32_t *entries,
                                 std::vector<uint32_t> &run_off, std::vector<uint32_t> &run_grp,
                                 std::vector<uint32_t> &prog)
{
    constexpr uint32_t kNever = 0xFFFFFFFFu;
    const unsigned n_groups = (n_surfaces + kAsmSurfaces - 1) / kAsmSurfaces;
    auto g_k0 = [&](unsigned g) { return offsets[g * kAsmSurfaces]; };
    auto g_k1 = [&](unsigned g) { return offsets[std::min(n_surfaces, (g + 1) * kAsmSurfaces)]; };
    // line -> the (at most two, a line may repeat) groups whose surfaces use it
    std::unordered_map<uint32_t, std::pair<uint32_t, uint32_t>> line_groups;
    line_groups.reserve(offsets[n_surfaces]);
    for (unsigned g = 0; g < n_groups; ++g)
        for (uint32_t k = g_k0(g); k < g_k1(g); ++k)
        {
            auto it = line_groups.emplace(entries[k] & kInvalidId, std::make_pair(g, kNever)).first;
            if (it->second.first != g && it->second.second == kNever)
                it->second.second = g;
        }
    // runs: from the first unassigned group, repeatedly the unassigned group sharing the most
    // lines with the run's last group (ties: the lowest index), up to kAsmGroups
    std::vector<char> assigned(n_groups, 0);
    std::vector<uint32_t> runs; // kAsmGroups entries per run
    std::unordered_map<uint32_t, unsigned> shared;
    for (unsigned g0 = 0; g0 < n_groups; ++g0)
    {
        if (assigned[g0])
            continue;
        uint32_t run[kAsmGroups];
        unsigned nr = 0;
        run[nr++] = g0;
        assigned[g0] = 1;
        while (nr < kAsmGroups)
        {
            shared.clear();
            const uint32_t last = run[nr - 1];
            for (uint32_t k = g_k0(last); k < g_k1(last); ++k)
            {
                const auto &lg = line_groups[entries[k] & kInvalidId];
                for (uint32_t g2 : {lg.first, lg.second})
                    if (g2 != kNever && !assigned[g2])
                        ++shared[g2];
            }
            uint32_t best = kNever;
            unsigned best_n = 0;
            for (const auto &kv : shared)
                if (kv.second > best_n || (kv.second == best_n && kv.first < best))
                    best = kv.first, best_n = kv.second;
            if (best == kNever)
                break;
            run[nr++] = best;
            assigned[best] = 1;
        }
        for (unsigned i = 0; i < kAsmGroups; ++i)
            runs.push_back(i < nr ? run[i] : kAsmNoGroup);
    }
    const unsigned n_runs = (unsigned)(runs.size() / kAsmGroups);
    run_grp = runs;
    run_off.assign(n_runs + 1, 0);
    prog.clear();
    prog.reserve(2 * ((size_t)offsets[n_surfaces] + n_surfaces + n_groups));
    std::vector<uint32_t> ent, next_use; // the run's entries (index into `entries`, or kNever: empty surface)
    std::unordered_map<uint32_t, uint32_t> last_seen;
    for (unsigned r = 0; r < n_runs; ++r)
    {
        ent.clear();
        for (unsigned i = 0; i < kAsmGroups; ++i)
        {
            const uint32_t g = runs[r * kAsmGroups + i];
            if (g == kAsmNoGroup)
                break;
            for (unsigned s = g * kAsmSurfaces; s < std::min(n_surfaces, (g + 1) * kAsmSurfaces); ++s)
                for (uint32_t k = offsets[s]; k < offsets[s + 1]; ++k)
                    ent.push_back(k);
        }
        // farthest-next-use slot choice over the run's entries (Belady)
        const uint32_t n = (uint32_t)ent.size();
        next_use.assign(n, kNever);
        last_seen.clear();
        for (uint32_t i = n; i-- > 0;)
        {
            const uint32_t l = entries[ent[i]] & kInvalidId;
            auto it = last_seen.find(l);
            next_use[i] = it == last_seen.end() ? kNever : it->second;
            last_seen[l] = i;
        }
        uint32_t slot_line[kAsmSlots], slot_next[kAsmSlots];
        for (unsigned q = 0; q < kAsmSlots; ++q)
            slot_line[q] = kNever, slot_next[q] = kNever;
        uint32_t i = 0;
        for (unsigned gi = 0; gi < kAsmGroups; ++gi)
        {
            const uint32_t g = runs[r * kAsmGroups + gi];
            if (g == kAsmNoGroup)
                break;
            const size_t group_start = prog.size();
            for (unsigned s = g * kAsmSurfaces; s < std::min(n_surfaces, (g + 1) * kAsmSurfaces); ++s)
            {
                if (offsets[s + 1] == offsets[s])
                {
                    prog.push_back(0);
                    prog.push_back(kAsmNone | kAsmEnd);
                    continue;
                }
                for (uint32_t k = offsets[s]; k < offsets[s + 1]; ++k, ++i)
                {
                    const uint32_t l = entries[k] & kInvalidId;
                    uint32_t op = kAsmEval;
                    int held = -1;
                    for (unsigned q = 0; q < kAsmSlots; ++q)
                        if (slot_line[q] == l)
                            held = (int)q;
                    if (held >= 0)
                    {
                        op = kAsmLoad | ((uint32_t)held << 2);
                        slot_next[held] = next_use[i];
                        if (next_use[i] == kNever)
                            slot_line[held] = kNever;
                    }
                    else if (next_use[i] != kNever)
                    {
                        // a free slot, else the one whose line is needed last (if later than this one)
                        int q_use = -1;
                        for (unsigned q = 0; q < kAsmSlots && q_use < 0; ++q)
                            if (slot_line[q] == kNever)
                                q_use = (int)q;
                        if (q_use < 0)
                        {
                            unsigned far = 0;
                            for (unsigned q = 1; q < kAsmSlots; ++q)
                                if (slot_next[q] > slot_next[far])
                                    far = q;
                            if (slot_next[far] > next_use[i])
                                q_use = (int)far;
                        }
                        if (q_use >= 0)
                        {
                            op = kAsmEvalStore | ((uint32_t)q_use << 2);
                            slot_line[q_use] = l;
                            slot_next[q_use] = next_use[i];
                        }
                    }
                    if (k + 1 == offsets[s + 1])
                        op |= kAsmEnd;
                    prog.push_back(entries[k]);
                    prog.push_back(op);
                }
            }
            if (((prog.size() - group_start) / 2) % 2)
            {
                prog.push_back(0); // padding: adds zero, ends no surface
                prog.push_back(kAsmNone);
            }
        }
        run_off[r + 1] = (uint32_t)(prog.size() / 2);
    }
}

extern "C" {

int pvl_mesh_create(pvl_mesh **out, unsigned n_points, unsigned n_lines, const uint32_t *lines, unsigned n_surfaces,
                    const uint32_t *surface_offsets, const uint32_t *surface_lines)
{
    PVL_API_LOCK;
    *out = nullptr;
    if (int st = ensure_init())
        return st;
    if (surface_offsets[0] != 0)
    {
        set_error("surface offsets must start at 0");
        return PVL_ERR_ARG;
    }
    for (unsigned i = 0; i < n_lines; ++i)
    {
        const uint32_t a = lines[2 * i] & kInvalidId, b = lines[2 * i + 1] & kInvalidId;
        if (a >= n_points || b >= n_points)
        {
            set_error("line %u references point (%u, %u) but the mesh has %u points", i, a, b, n_points);
            return PVL_ERR_ARG;
        }
    }
    const uint32_t n_entries = surface_offsets[n_surfaces];
    unsigned max_len = 0;
    for (unsigned s = 0; s < n_surfaces; ++s)
    {
        if (surface_offsets[s + 1] < surface_offsets[s])
        {
            set_error("surface offsets are not monotone at surface %u", s);
            return PVL_ERR_ARG;
        }
        const unsigned len = surface_offsets[s + 1] - surface_offsets[s];
        if (len > max_len)
            max_len = len;
    }
    for (uint32_t k = 0; k < n_entries; ++k)
    {
        if ((surface_lines[k] & kInvalidId) >= n_lines)
        {
            set_error("surface entry %u references line %u but the mesh has %u lines", k,
                      surface_lines[k] & kInvalidId, n_lines);
            return PVL_ERR_ARG;
        }
    }
    // line -> (surface | sign) transpose, surfaces in increasing order inside a line
    std::vector<uint32_t> loff(n_lines + 1, 0), lsurf(n_entries);
    for (uint32_t k = 0; k < n_entries; ++k)
        loff[(surface_lines[k] & kInvalidId) + 1] += 1;
    for (unsigned i = 0; i < n_lines; ++i)
        loff[i + 1] += loff[i];
    {
        std::vector<uint32_t> fill(loff.begin(), loff.end() - 1);
        for (unsigned s = 0; s < n_surfaces; ++s)
            for (uint32_t k = surface_offsets[s]; k < surface_offsets[s + 1]; ++k)
            {
                const uint32_t l = surface_lines[k] & kInvalidId;
                lsurf[fill[l]++] = s | (surface_lines[k] & kReversed);
            }
    }
    pvl_mesh *m = new pvl_mesh;
    m->n_points = n_points;
    m->n_lines = n_lines;
    m->n_surfaces = n_surfaces;
    m->max_surface_lines = max_len;
    auto up = [&](uint32_t **dst, const uint32_t *src, size_t count) -> int {
        const size_t bytes = (count ? count : 1) * sizeof(uint32_t);
        if (hipMalloc(dst, bytes) != hipSuccess)
        {
            set_error("device allocation of %zu bytes for the mesh failed", bytes);
            return PVL_ERR_NOMEM;
        }
        if (count)
            return check_hip(hipMemcpy(*dst, src, count * sizeof(uint32_t), hipMemcpyHostToDevice), "hipMemcpy");
        return PVL_OK;
    };
    int st = up(&m->d_lines, lines, 2 * (size_t)n_lines);
    if (!st)
        st = up(&m->d_offsets, surface_offsets, n_surfaces + 1);
    if (!st)
        st = up(&m->d_entries, surface_lines, n_entries);
    if (!st)
        st = up(&m->d_line_offsets, loff.data(), n_lines + 1);
    if (!st)
        st = up(&m->d_line_surfs, lsurf.data(), n_entries);
    if (!st)
    {
        std::vector<uint32_t> run_off, run_grp, prog;
        build_assembly_program(n_surfaces, surface_offsets, surface_lines, run_off, run_grp, prog);
        m->n_asm_runs = (unsigned)(run_off.size() - 1);
        st = up(&m->d_asm_off, run_off.data(), run_off.size());
        if (!st)
            st = up(&m->d_asm_grp, run_grp.data(), run_grp.size());
        if (!st)
            st = up(&m->d_asm, prog.data(), prog.size());
    }
    if (st)
    {
        pvl_mesh_destroy(m);
        return st;
    }
    *out = m;
    return PVL_OK;
}

int pvl_mesh_destroy(pvl_mesh *m)
{
    PVL_API_LOCK;
    if (!m)
        return PVL_OK;
    (void)hipStreamSynchronize(stream());
    (void)hipFree(m->d_lines);
    (void)hipFree(m->d_offsets);
    (void)hipFree(m->d_entries);
    (void)hipFree(m->d_line_offsets);
    (void)hipFree(m->d_line_surfs);
    (void)hipFree(m->d_asm_off);
    (void)hipFree(m->d_asm_grp);
    (void)hipFree(m->d_asm);
    delete m;
    return PVL_OK;
}

int pvl_mesh_counts(const pvl_mesh *m, unsigned *n_points, unsigned *n_lines, unsigned *n_surfaces)
{
    PVL_API_LOCK;
    *n_points = m->n_points;
    *n_lines = m->n_lines;
    *n_surfaces = m->n_surfaces;
    return PVL_OK;
}

} // extern "C"
