"""Pins of the oracle (oracle/pvl_oracle.c) before it is trusted as the checker.

Three independent anchors:
1. the reference's own C core compiled unmodified into oracle/_ref (bit-for-bit equality);
2. the golden fixtures of the reference's C tests (test/from_points, test/dual);
3. closed forms and invariants of the reference kernel (perpendicular bisector, cut-off
   zeros, -2*pi self terms of planar convex panels, coplanar decoupling, segment reversal,
   closed-mesh row sums, |K| <= pi, K perpendicular to the segment).
"""

import numpy as np
import pytest

from oracle import oracle as O
from pyvl_amd import geometry as G

GOLD = __import__("pathlib").Path(__file__).parent / "golden" / "reference_fixtures"
needs_ref = pytest.mark.skipif(O.load_ref() is None, reason="reference sources not built into oracle/_ref")


def _random_mesh(seed, n_span=7, n_chord=5, warp=0.3):
    rng = np.random.default_rng(seed)
    pos, el, _ = G.flat_wing(n_span, n_chord, 3.0, 1.0)
    pos = pos + rng.normal(0.0, 0.05, pos.shape)
    pos[:, 2] += warp * np.sin(pos[:, 0] * 2.0) * np.cos(pos[:, 1])
    return pos, el


def _cases():
    pos, el, _ = G.flat_wing(16, 4, 8.0, 1.0)
    yield "wing16x4", pos, el
    pos, el = G.uv_sphere(16, 8, sigma=1e-3, seed=3)
    yield "sphere16x8", pos, el
    pos, el = _random_mesh(11)
    yield "warped", pos, el
    pos, el = G.ellipsoid(12, 10, sigma_rel=1e-3, seed=5)
    yield "ellipsoid12x10", pos, el


CASES = list(_cases())


@needs_ref
@pytest.mark.parametrize("name,pos,el", CASES, ids=[c[0] for c in CASES])
def test_topology_bit_exact_vs_reference(name, pos, el):
    t = O.mesh_from_elements(pos.shape[0], el)
    r = O.ref_mesh_from_elements(pos.shape[0], el)
    assert np.array_equal(t.lines, r.lines)
    assert np.array_equal(t.offsets, r.offsets)
    assert np.array_equal(t.entries, r.entries)
    d = O.mesh_dual(t)
    dr = O.ref_mesh_dual(t)
    for a, b in zip(d, dr):
        assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("name,pos,el", CASES, ids=[c[0] for c in CASES])
def test_induction_bit_exact_vs_reference(name, pos, el):
    t = O.mesh_from_elements(pos.shape[0], el)
    n = O.surface_normal(t, pos)
    c = O.surface_center(t, pos)
    assert np.array_equal(n, O.ref_surface_normal(t, pos))
    assert np.array_equal(c, O.ref_surface_center(t, pos))
    rng = np.random.default_rng(1)
    pts = np.concatenate([c, pos[:5], rng.normal(0, 1.5, (7, 3))])
    assert np.array_equal(O.line_induction(t, pos, pts, 1e-6), O.ref_line_induction(t, pos, pts, 1e-6))
    nn = rng.normal(size=pts.shape)
    nn /= np.linalg.norm(nn, axis=1)[:, None]
    assert np.array_equal(O.induction_matrix3(t, pos, pts, nn, 1e-6), O.ref_induction_matrix3(t, pos, pts, nn, 1e-6))
    assert np.array_equal(O.induction_matrix(t, pos, pts, 1e-6), O.ref_induction_matrix(t, pos, pts, 1e-6))


def test_golden_from_points():
    """test/from_points: mesh_from_elements(test_points.dat) must equal test_mesh.msh."""
    rows = [list(map(int, ln.split())) for ln in (GOLD / "from_points_test_points.dat").read_text().splitlines()
            if ln.strip()]
    _, want = O.read_msh(GOLD / "from_points_test_mesh.msh")
    got = O.mesh_from_elements(want.n_points, rows)
    assert np.array_equal(got.lines, want.lines)
    assert np.array_equal(got.offsets, want.offsets)
    assert np.array_equal(got.entries, want.entries)


def test_golden_dual():
    """test/dual: mesh_dual_from_primal(test_mesh.msh) must equal test_dual.msh."""
    _, prim = O.read_msh(GOLD / "dual_test_mesh.msh")
    _, want = O.read_msh(GOLD / "dual_test_dual.msh")
    dl, do, de = O.mesh_dual(prim)
    assert np.array_equal(dl, want.lines)
    assert np.array_equal(do, want.offsets)
    assert np.array_equal(de, want.entries)


def _single_line(r1, r2):
    pos = np.array([r1, r2], np.float64)
    t = O.Topo(2, np.array([[0, 1]], np.uint32), np.array([0], np.uint32), np.zeros(0, np.uint32))
    return t, pos


@pytest.mark.parametrize("a,h", [(0.5, 0.3), (2.0, 0.01), (1e-3, 5.0), (1.0, 1.0)])
def test_closed_form_perpendicular_bisector(a, h):
    t, pos = _single_line((-a, 0, 0), (a, 0, 0))
    k = O.line_induction(t, pos, np.array([[0.0, h, 0.0]]), 1e-6)[0, 0]
    assert k[0] == 0.0 and k[1] == 0.0
    # h^2 = |dr|^2 - t^2 (src/core/flow_solver.c:120-123) loses ~log10((a/h)^2) digits
    rtol = 1e-14 + 4e-16 * (a * a + h * h) / (h * h)
    assert k[2] == pytest.approx(2.0 * np.arctan(a / h), rel=rtol)


def test_cutoff_zeros_on_endpoints_and_line():
    """test/pytests/test_mesh.py:105-111 (control point on a segment end gives exact 0),
    plus points on the segment's line and a segment shorter than tol."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float64)
    t = O.mesh_from_elements(4, [[0, 1, 2, 3]])
    m = O.line_induction(t, pos, pos, 1e-6)
    for i in range(4):
        assert np.all(m[i, i, :] == 0.0)
    t1, p1 = _single_line((0, 0, 0), (1, 0, 0))
    on_line = np.array([[-0.5, 0, 0], [2.0, 0, 0], [0.3, 0, 0]])
    assert np.all(O.line_induction(t1, p1, on_line, 1e-6) == 0.0)
    t2, p2 = _single_line((0, 0, 0), (1e-7, 0, 0))
    assert np.all(O.line_induction(t2, p2, np.array([[0.0, 1.0, 0.0]]), 1e-6) == 0.0)


def test_reversal_translation_and_bounds():
    rng = np.random.default_rng(7)
    r1, r2 = rng.normal(size=3), rng.normal(size=3)
    pts = rng.normal(0, 2, (50, 3))
    ta, pa = _single_line(r1, r2)
    tb, pb = _single_line(r2, r1)
    ka = O.line_induction(ta, pa, pts, 1e-9)[:, 0]
    kb = O.line_induction(tb, pb, pts, 1e-9)[:, 0]
    np.testing.assert_allclose(kb, -ka, rtol=1e-10, atol=1e-13)
    shift = np.array([0.3, -1.2, 0.7])
    kc = O.line_induction(ta, pa + shift, pts + shift, 1e-9)[:, 0]
    np.testing.assert_allclose(kc, ka, rtol=1e-9, atol=1e-12)
    assert np.all(np.linalg.norm(ka, axis=1) <= np.pi * (1 + 1e-14))
    e = (r2 - r1) / np.linalg.norm(r2 - r1)
    assert np.max(np.abs(ka @ e)) < 1e-13
    # |K| equals the subtended angle
    ang = np.array([np.arccos(np.clip(np.dot(p - r1, p - r2) / np.linalg.norm(p - r1) / np.linalg.norm(p - r2),
                                      -1, 1)) for p in pts])
    np.testing.assert_allclose(np.linalg.norm(ka, axis=1), ang, rtol=1e-9, atol=1e-12)


def test_planar_self_term_and_coplanar_decoupling():
    """A_ii = -2*pi for a planar convex panel at its vertex centroid; A_ij ~ 0 for coplanar
    exterior panels (the decoupling follows from the normal convention of
    src/core/mesh.c:48-68)."""
    rng = np.random.default_rng(2)
    for _ in range(10):
        # random convex quad in a random plane
        ang = np.sort(rng.uniform(0, 2 * np.pi, 4))
        rad = rng.uniform(0.5, 1.5, 4)
        q = np.stack([rad * np.cos(ang), rad * np.sin(ang), np.zeros(4)], axis=1)
        q2 = q + np.array([5.0, 0.3, 0.0])
        basis, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        pos = np.concatenate([q, q2]) @ basis.T + rng.normal(size=3)
        t = O.mesh_from_elements(8, [[0, 1, 2, 3], [4, 5, 6, 7]])
        n = O.surface_normal(t, pos)
        c = O.surface_center(t, pos)
        a = O.induction_matrix3(t, pos, c, n, 1e-6)
        assert a[0, 0] == pytest.approx(-2 * np.pi, rel=1e-13)
        assert a[1, 1] == pytest.approx(-2 * np.pi, rel=1e-13)
        assert abs(a[0, 1]) < 1e-13 and abs(a[1, 0]) < 1e-13


def test_closed_mesh_row_sums_vanish():
    pos, el = G.uv_sphere(24, 12, sigma=1e-3, seed=9)
    t = O.mesh_from_elements(pos.shape[0], el)
    n = O.surface_normal(t, pos)
    c = O.surface_center(t, pos)
    a = O.induction_matrix3(t, pos, c, n, 1e-6)
    assert np.max(np.abs(a.sum(axis=1))) < 1e-12 * np.max(np.abs(a))


def test_reference_mesh_expectations():
    """test/pytests/test_mesh.py:18-68: unit normals, perpendicular to triangle edges;
    centres equal to vertex means."""
    positions = np.array([[0, 2, 0.5], [0.4, 2, 0.4], [4, 0.1, 0.2], [3, -0.1, 0.2], [2, 1, 0.2], [2, -1, 3]])
    elements = [[0, 1, 2, 3], [2, 3, 4], [0, 1, 4], [2, 4, 5]]
    t = O.mesh_from_elements(6, elements)
    n = O.surface_normal(t, positions)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0)
    for v, e in zip(n, elements):
        if len(e) != 3:
            continue
        for i in range(3):
            assert abs(np.dot(positions[e[i]] - positions[e[i - 1]], v)) < 1e-12
    c = O.surface_center(t, positions)
    for v, e in zip(c, elements):
        assert np.allclose(v, np.mean(positions[e], axis=0), rtol=1e-15, atol=1e-15)


def test_point_velocity_equals_tensor_contraction():
    pos, el = _random_mesh(4)
    t = O.mesh_from_elements(pos.shape[0], el)
    rng = np.random.default_rng(0)
    circ = rng.normal(size=t.n_surfaces)
    pts = rng.normal(0, 1.0, (20, 3)) + np.array([0.5, 0, 0.5])
    v = O.point_velocity(t, pos, circ, pts, 1e-6)
    ten = O.induction_matrix(t, pos, pts, 1e-6)
    want = np.einsum("msc,s->mc", ten, circ)
    np.testing.assert_allclose(v, want, rtol=1e-12, atol=1e-13)


def test_small_lu_matches_lapack():
    sla = pytest.importorskip("scipy.linalg")
    rng = np.random.default_rng(3)
    a = rng.normal(size=(9, 9))
    lu, piv, info = O.lu_factor_small(a)
    assert info == 0
    lu_r, piv_r = sla.lu_factor(a)
    assert np.array_equal(piv, piv_r)
    np.testing.assert_allclose(lu, lu_r, rtol=1e-12, atol=1e-14)
