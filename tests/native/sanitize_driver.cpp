// Host-code sanitizer run (AddressSanitizer + UndefinedBehaviorSanitizer), CPU only.
//
// Drives the library's host topology builders (pvl_mesh_from_elements, pvl_mesh_dual,
// pvl_wake_topology and the assembly program builder in pyvl_amd/csrc/topology.hip; the
// program is run on the CPU against the oracle's induction_matrix3) and the oracle's C restatement
// (oracle/pvl_oracle.c: mesh construction, dual, line induction, surface reductions, point
// velocity, line forces, LU) over seeded random meshes -- quads, triangles, mixed and
// degenerate elements, shared and boundary edges -- and checks that both topologies agree.
// Built by `make -C pyvl_amd/csrc sanitize`; no GPU call is made.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

extern "C" {
int pvl_mesh_from_elements(unsigned, const unsigned *, const unsigned *, unsigned *, uint32_t *, uint32_t *,
                           uint32_t *);
int pvl_mesh_dual(unsigned, unsigned, const uint32_t *, unsigned, const uint32_t *, const uint32_t *, uint32_t *,
                  uint32_t *, uint32_t *);
int pvl_wake_topology(unsigned, unsigned, unsigned *, uint32_t *, uint32_t *, uint32_t *);
int orc_mesh_from_elements(unsigned, const unsigned *, const unsigned *, unsigned *, uint32_t *, uint32_t *,
                           uint32_t *);
int orc_mesh_dual(unsigned, unsigned, const uint32_t *, unsigned, const uint32_t *, const uint32_t *, uint32_t *,
                  uint32_t *, uint32_t *);
void orc_line_induction(unsigned, const uint32_t *, const double *, unsigned, const double *, double *, double);
void orc_induction_matrix3(unsigned, const uint32_t *, unsigned, const uint32_t *, const uint32_t *, const double *,
                           unsigned, const double *, const double *, double *, double);
void orc_point_velocity(unsigned, const uint32_t *, unsigned, const uint32_t *, const uint32_t *, const double *,
                        const double *, unsigned, const double *, double *, double);
void orc_line_forces(unsigned, const uint32_t *, const uint32_t *, const double *, const double *, const double *,
                     double *);
int orc_lu_factor(unsigned, double *, int *);
}

// The assembly program builder (pyvl_amd/csrc/topology.hip; operations as in pvl_internal.hpp:
// kind in bits 0-1 -- 0 evaluate, 1 evaluate and store, 2 load, 3 none --, slot in bits 2-5,
// bit 8 on a surface's last entry; a none entry without bit 8 pads a group to an even length).
namespace pvl
{
void build_assembly_program(unsigned n_surfaces, const uint32_t *offsets, const uint32_t *entries,
                            std::vector<uint32_t> &run_off, std::vector<uint32_t> &run_grp,
                            std::vector<uint32_t> &prog);
}

constexpr unsigned kAsmSurfaces = 8; // surfaces per group, as pvl::kAsmSurfaces
constexpr unsigned kAsmGroups = 4;   // groups per run, as pvl::kAsmGroups
constexpr unsigned kAsmSlots = 10;   // LDS slots per lane, as pvl::kAsmSlots

// Runs the program as the kernel does, for one control point, with the per-line values
// v[line] (n . K of that line at the point): slots hold (line, value); a load must find its
// own line; every group starts on an even entry and every surface is produced exactly once.
// Returns the per-surface sums, or an empty vector when the program is malformed.
static std::vector<double> run_program(unsigned n_el, const std::vector<uint32_t> &o, const std::vector<uint32_t> &e,
                                       const std::vector<uint32_t> &run_off, const std::vector<uint32_t> &run_grp,
                                       const std::vector<uint32_t> &prog, const std::vector<double> &v)
{
    const unsigned n_runs = (unsigned)run_off.size() - 1;
    if (run_grp.size() != (size_t)n_runs * kAsmGroups || run_off[0] != 0)
        return {};
    std::vector<double> out(n_el, 0.0);
    std::vector<char> done(n_el, 0);
    for (unsigned r = 0; r < n_runs; ++r)
    {
        uint32_t slot_line[kAsmSlots];
        double slot_val[kAsmSlots];
        for (unsigned q = 0; q < kAsmSlots; ++q)
            slot_line[q] = 0xFFFFFFFFu, slot_val[q] = 0.0;
        double acc = 0.0;
        uint32_t k = run_off[r];
        for (unsigned gi = 0; gi < kAsmGroups; ++gi)
        {
            const uint32_t g = run_grp[r * kAsmGroups + gi];
            if (g == 0xFFFFFFFFu)
                break;
            if ((k - run_off[r]) % 2 != 0 || (size_t)g * kAsmSurfaces >= n_el)
                return {};
            const unsigned s1 = std::min(n_el, (g + 1) * kAsmSurfaces);
            for (unsigned s = g * kAsmSurfaces; s < s1; ++s)
            {
                if (done[s])
                    return {};
                uint32_t k_csr = o[s];
                for (;;)
                {
                    if (k >= run_off[r + 1])
                        return {};
                    const uint32_t word = prog[2 * k], op = prog[2 * k + 1];
                    ++k;
                    const uint32_t kind = op & 3u, slot = (op >> 2) & 15u, line = word & 0x7FFFFFFFu;
                    if (kind != 3 && slot >= kAsmSlots && kind != 0)
                        return {};
                    double val = 0.0;
                    if (kind == 3)
                    {
                        if ((op & (1u << 8)) && o[s] != o[s + 1])
                            return {};
                    }
                    else
                    {
                        if (k_csr >= o[s + 1] || word != e[k_csr++])
                            return {};
                        if (kind == 2)
                        {
                            if (slot_line[slot] != line)
                                return {};
                            val = slot_val[slot];
                        }
                        else
                        {
                            val = v[line];
                            if (kind == 1)
                                slot_line[slot] = line, slot_val[slot] = val;
                        }
                    }
                    acc = (word & 0x80000000u) ? acc - val : acc + val;
                    if (op & (1u << 8))
                    {
                        if (k_csr != o[s + 1])
                            return {};
                        out[s] = acc;
                        done[s] = 1;
                        acc = 0.0;
                        break;
                    }
                }
            }
            // a padding entry (none, no end) completes an odd group
            if ((k - run_off[r]) % 2 != 0)
            {
                if (k >= run_off[r + 1] || prog[2 * k + 1] != 3u)
                    return {};
                ++k;
            }
        }
        if (k != run_off[r + 1])
            return {};
    }
    for (unsigned s = 0; s < n_el; ++s)
        if (!done[s])
            return {};
    return out;
}

static int fail(const char *what, int k)
{
    std::fprintf(stderr, "sanitize: mismatch in %s (case %d)\n", what, k);
    return 1;
}

int main()
{
    std::mt19937 rng(12345);
    int bad = 0;
    for (int k = 0; k < 60; ++k)
    {
        // a structured grid of quads, some split into triangles, plus random extra elements
        const unsigned nx = 2 + rng() % 9, ny = 2 + rng() % 7;
        const unsigned n_points = (nx + 1) * (ny + 1) + 3;
        std::vector<unsigned> counts, flat;
        for (unsigned j = 0; j < ny; ++j)
            for (unsigned i = 0; i < nx; ++i)
            {
                const unsigned a = j * (nx + 1) + i, b = a + 1, c = b + nx + 1, d = a + nx + 1;
                if (rng() % 4 == 0)
                {
                    counts.push_back(3), flat.insert(flat.end(), {a, b, c});
                    counts.push_back(3), flat.insert(flat.end(), {a, c, d});
                }
                else
                    counts.push_back(4), flat.insert(flat.end(), {a, b, c, d});
            }
        for (int e = 0; e < 3; ++e) // random elements, possibly repeating points
        {
            const unsigned n = 2 + rng() % 4;
            counts.push_back(n);
            for (unsigned q = 0; q < n; ++q)
                flat.push_back(rng() % n_points);
        }
        if (k % 7 == 0)
            counts.push_back(0);
        const unsigned n_el = (unsigned)counts.size(), total = (unsigned)flat.size();
        std::vector<uint32_t> l1(2 * total), o1(n_el + 1), e1(total), l2(2 * total), o2(n_el + 1), e2(total);
        unsigned nl1 = 0, nl2 = 0;
        if (pvl_mesh_from_elements(n_el, counts.data(), flat.data(), &nl1, l1.data(), o1.data(), e1.data()) != 0)
            return fail("pvl_mesh_from_elements status", k);
        orc_mesh_from_elements(n_el, counts.data(), flat.data(), &nl2, l2.data(), o2.data(), e2.data());
        if (nl1 != nl2)
            bad |= fail("line count", k);
        for (unsigned i = 0; i < 2 * nl1 && !bad; ++i)
            bad |= (l1[i] != l2[i]) ? fail("lines", k) : 0;
        for (unsigned i = 0; i <= n_el && !bad; ++i)
            bad |= (o1[i] != o2[i]) ? fail("offsets", k) : 0;
        for (unsigned i = 0; i < total && !bad; ++i)
            bad |= (e1[i] != e2[i]) ? fail("entries", k) : 0;
        std::vector<uint32_t> dl1(2 * nl1 + 2), do1(n_points + 1), de1(2 * nl1 + 2), dl2(2 * nl1 + 2),
            do2(n_points + 1), de2(2 * nl1 + 2);
        if (pvl_mesh_dual(n_points, nl1, l1.data(), n_el, o1.data(), e1.data(), dl1.data(), do1.data(),
                          de1.data()) != 0)
            return fail("pvl_mesh_dual status", k);
        orc_mesh_dual(n_points, nl1, l1.data(), n_el, o1.data(), e1.data(), dl2.data(), do2.data(), de2.data());
        for (unsigned i = 0; i < 2 * nl1 && !bad; ++i)
            bad |= (dl1[i] != dl2[i]) ? fail("dual lines", k) : 0;
        for (unsigned i = 0; i <= n_points && !bad; ++i)
            bad |= (do1[i] != do2[i]) ? fail("dual offsets", k) : 0;
        for (unsigned i = 0; i < do1[n_points] && !bad; ++i)
            bad |= (de1[i] != de2[i]) ? fail("dual entries", k) : 0;

        // oracle numerics on the same mesh
        std::normal_distribution<double> nd(0.0, 1.0);
        std::vector<double> pos(3 * n_points), circ(n_el), vel(3 * n_points);
        for (auto &v : pos)
            v = nd(rng);
        for (auto &v : circ)
            v = nd(rng);
        for (auto &v : vel)
            v = nd(rng);
        const unsigned n_cpts = 7;
        std::vector<double> cpts(3 * n_cpts), nrm(3 * n_cpts, 0.0);
        for (auto &v : cpts)
            v = nd(rng);
        for (unsigned i = 0; i < n_cpts; ++i)
            nrm[3 * i + 2] = 1.0;
        std::vector<double> lk(3 * (size_t)n_cpts * (nl1 ? nl1 : 1)), a3((size_t)n_cpts * n_el), pv(3 * n_cpts),
            lf(3 * (size_t)(nl1 ? nl1 : 1));
        orc_line_induction(nl1, l1.data(), pos.data(), n_cpts, cpts.data(), lk.data(), 1e-6);
        orc_induction_matrix3(nl1, l1.data(), n_el, o1.data(), e1.data(), pos.data(), n_cpts, cpts.data(),
                              nrm.data(), a3.data(), 1e-6);
        orc_point_velocity(nl1, l1.data(), n_el, o1.data(), e1.data(), pos.data(), circ.data(), n_cpts, cpts.data(),
                           pv.data(), 1e-6);
        orc_line_forces(nl1, l1.data(), dl1.data(), circ.data(), pos.data(), vel.data(), lf.data());
        // the assembly program reproduces every surface's sum for every control point
        {
            std::vector<uint32_t> run_off, run_grp, prog;
            pvl::build_assembly_program(n_el, o1.data(), e1.data(), run_off, run_grp, prog);
            for (unsigned c = 0; c < n_cpts && !bad; ++c)
            {
                std::vector<double> v(nl1 ? nl1 : 1);
                for (unsigned l = 0; l < nl1; ++l)
                {
                    const double *kv = &lk[3 * ((size_t)c * nl1 + l)];
                    v[l] = kv[0] * nrm[3 * c] + kv[1] * nrm[3 * c + 1] + kv[2] * nrm[3 * c + 2];
                }
                const std::vector<double> got = run_program(n_el, o1, e1, run_off, run_grp, prog, v);
                if (got.size() != n_el)
                {
                    bad |= fail("assembly program structure", k);
                    break;
                }
                for (unsigned s = 0; s < n_el && !bad; ++s)
                {
                    const double want = a3[(size_t)c * n_el + s];
                    if (std::fabs(got[s] - want) > 1e-12 * (1.0 + std::fabs(want)))
                        bad |= fail("assembly program sum", k);
                }
            }
        }
        const unsigned n = 5 + k % 20;
        std::vector<double> a((size_t)n * n);
        std::vector<int> piv(n);
        for (auto &v : a)
            v = nd(rng);
        orc_lu_factor(n, a.data(), piv.data());
    }
    for (unsigned strips = 1; strips < 6 && !bad; ++strips)
        for (unsigned rows = 1; rows < 5; ++rows)
        {
            const unsigned nl = strips * (3 * rows + 1);
            std::vector<uint32_t> l(2 * nl), o(strips * rows + 1), e(4 * strips * rows);
            unsigned got = 0;
            if (pvl_wake_topology(strips, rows, &got, l.data(), o.data(), e.data()) != 0 || got != nl)
                return fail("pvl_wake_topology", (int)(strips * 10 + rows));
        }
    if (!bad)
        std::printf("sanitize: 60 meshes, topology equal to the oracle, no sanitizer report\n");
    return bad;
}
