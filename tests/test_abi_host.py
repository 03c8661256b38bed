"""CPU-side checks of libpvl_hip.so: it loads, exports every symbol include/pvl.h declares,
and its host-only topology code matches the oracle bit for bit.  The Python ``Mesh``
mirror's argument validation follows the reference (src/common.h:8-66,
src/meshobject.c:25-124, :384-473).  No device work is done here.
"""

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
from pyvl_amd import _lib
from pyvl_amd import geometry as G
from pyvl_amd.cvl import INVALID_ID, Mesh

ROOT = Path(__file__).resolve().parents[1]
GOLD = ROOT / "tests" / "golden" / "reference_fixtures"


def _header_symbols():
    txt = "\n".join((ROOT / "include").glob("*.h").__next__().read_text().splitlines())
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pvl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    syms = _header_symbols()
    assert len(syms) > 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python binding declares a signature for each one
    assert sorted(_lib.SIGNATURES) == syms


def _host_from_elements(elements):
    counts = np.array([len(e) for e in elements], np.uint32)
    flat = np.concatenate([np.asarray(e, np.uint32) for e in elements])
    lines = np.empty((flat.size, 2), np.uint32)
    off = np.empty(counts.size + 1, np.uint32)
    ent = np.empty(flat.size, np.uint32)
    nl = C.c_uint()
    _lib.call("pvl_mesh_from_elements", counts.size, counts.ctypes.data_as(C.POINTER(C.c_uint)),
              flat.ctypes.data_as(C.POINTER(C.c_uint)), C.byref(nl), lines.ctypes.data_as(C.POINTER(C.c_uint32)),
              off.ctypes.data_as(C.POINTER(C.c_uint32)), ent.ctypes.data_as(C.POINTER(C.c_uint32)))
    return lines[: nl.value], off, ent


def _random_elements(seed, n_pts=40, n_el=60):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_el):
        k = int(rng.integers(3, 6))
        out.append(list(map(int, rng.choice(n_pts, size=k, replace=False))))
    return n_pts, out


@pytest.mark.parametrize("seed", range(6))
def test_host_from_elements_matches_oracle(seed):
    n_pts, el = _random_elements(seed)
    lines, off, ent = _host_from_elements(el)
    t = O.mesh_from_elements(n_pts, el)
    assert np.array_equal(lines, t.lines)
    assert np.array_equal(off, t.offsets)
    assert np.array_equal(ent, t.entries)


def test_host_from_elements_golden():
    rows = [list(map(int, ln.split())) for ln in (GOLD / "from_points_test_points.dat").read_text().splitlines()
            if ln.strip()]
    _, want = O.read_msh(GOLD / "from_points_test_mesh.msh")
    lines, off, ent = _host_from_elements(rows)
    assert np.array_equal(lines, want.lines)
    assert np.array_equal(off, want.offsets)
    assert np.array_equal(ent, want.entries)


@pytest.mark.parametrize("seed", range(4))
def test_host_dual_matches_oracle(seed):
    n_pts, el = _random_elements(seed + 100)
    m = Mesh(n_pts, el)
    d = m.compute_dual()
    t = O.mesh_from_elements(n_pts, el)
    dl, do, de = O.mesh_dual(t)
    assert np.array_equal(d.line_data, dl)
    assert np.array_equal(d.surface_offsets, do)
    assert np.array_equal(d.surface_lines, de)


def test_host_dual_golden():
    pos, prim = O.read_msh(GOLD / "dual_test_mesh.msh")
    _, want = O.read_msh(GOLD / "dual_test_dual.msh")
    m = Mesh._from_arrays(prim.n_points, prim.lines, prim.offsets, prim.entries)
    d = m.compute_dual()
    assert np.array_equal(d.line_data, want.lines)
    assert np.array_equal(d.surface_offsets, want.offsets)
    assert np.array_equal(d.surface_lines, want.entries)


@pytest.mark.parametrize("n_strips,rows", [(1, 1), (1, 5), (3, 4), (64, 7)])
def test_wake_topology_closed_form(n_strips, rows):
    """pvl_wake_topology == Mesh(2(R+1)n, rings) of python/pyvl/wake_models.py:35-52."""
    R = rows
    npr = R + 1
    ring = np.empty((n_strips, R, 4), np.uint32)
    ring[:, :, 0] = np.arange(R)
    ring[:, :, 1] = np.arange(R) + npr
    ring[:, :, 2] = np.arange(R) + npr + 1
    ring[:, :, 3] = np.arange(R) + 1
    ring += (2 * npr * np.arange(n_strips, dtype=np.uint32))[:, None, None]
    el = ring.reshape(-1, 4).tolist()
    t = O.mesh_from_elements(2 * npr * n_strips, el)
    w = Mesh.wake_rings(n_strips, R)
    assert w.n_points == 2 * npr * n_strips
    assert w.n_lines == n_strips * (3 * R + 1)
    assert np.array_equal(w.line_data, t.lines)
    assert np.array_equal(w.surface_offsets, t.offsets)
    assert np.array_equal(w.surface_lines, t.entries)


# --- reference pytests on the Mesh mirror (test/pytests/test_mesh.py:8-102) -------------
def test_mesh_construction():
    elements = [[0, 1, 2, 3], [2, 3, 4], [0, 1, 4], [2, 4, 5]]
    msh = Mesh(max(max(e) for e in elements), elements)  # index == n_points accepted
    flat_elements = np.array(sum(elements, start=[]))
    element_count, connectivity = msh.to_element_connectivity()
    assert np.all(element_count == [len(e) for e in elements])
    assert all(connectivity == flat_elements)


def test_mesh_merge():
    msh1 = Mesh(6, [[0, 1, 2, 3], [2, 3, 4], [0, 1, 4], [2, 4, 5]])
    msh2 = Mesh(6, [[2, 1, 0], [1, 2, 3], [4, 1, 5]])
    msh3 = Mesh(6, [[4, 1, 0], [5, 2, 4]])
    merged = Mesh.merge_meshes(msh1, msh2, msh3)
    assert merged.n_surfaces == msh1.n_surfaces + msh2.n_surfaces + msh3.n_surfaces
    assert merged.n_lines == msh1.n_lines + msh2.n_lines + msh3.n_lines
    assert merged.n_points == msh1.n_points + msh2.n_points + msh3.n_points
    n_lines = n_surfaces = point_offset = 0
    for m in (msh1, msh2, msh3):
        for i_s in range(m.n_surfaces):
            sm, so = merged.get_surface(n_surfaces), m.get_surface(i_s)
            for l1, l2 in zip(sm.lines, so.lines, strict=True):
                assert l1.begin == l2.begin + point_offset
                assert l1.end == l2.end + point_offset
            n_surfaces += 1
        for i_l in range(m.n_lines):
            lm, lo = merged.get_line(n_lines), m.get_line(i_l)
            assert lm.begin == lo.begin + point_offset
            assert lm.end == lo.end + point_offset
            n_lines += 1
        point_offset += m.n_points


def test_mesh_validation_mirrors_reference():
    with pytest.raises(ValueError):
        Mesh(4, [[0, 1]])
    with pytest.raises(ValueError):
        Mesh(3, [[0, 1, 4]])
    with pytest.raises(TypeError):
        Mesh(3, [[0, 1, 2], 5])
    pos, el, _ = G.flat_wing(4, 2, 4.0, 1.0)
    m = Mesh(pos.shape[0], el)
    cp = np.zeros((3, 3))
    with pytest.raises(TypeError):
        m.induction_matrix3(1e-6, pos, cp)  # 3 args
    with pytest.raises(TypeError):
        m.induction_matrix3("x", pos, cp, cp)
    with pytest.raises(TypeError):
        m.induction_matrix3(1e-6, pos.tolist(), cp, cp)
    with pytest.raises(ValueError):
        m.induction_matrix3(1e-6, pos[:-1], cp, cp)  # wrong point count
    with pytest.raises(ValueError):
        m.induction_matrix3(1e-6, pos.astype(np.float32), cp, cp)
    with pytest.raises(ValueError):
        m.induction_matrix3(1e-6, np.asfortranarray(np.zeros((pos.shape[0], 3)))[:, :], cp, cp[:2])
    with pytest.raises(ValueError):
        m.induction_matrix3(1e-6, pos, cp, cp, np.zeros((3, m.n_surfaces + 1)))
    ro = np.zeros((3, m.n_surfaces))
    ro.flags.writeable = False
    with pytest.raises(ValueError):
        m.induction_matrix3(1e-6, pos, cp, cp, ro)
    with pytest.raises(ValueError):
        m.induction_matrix3(1e-6, pos, cp, cp, None, np.zeros(5))  # line buffer too small
    with pytest.raises(TypeError):
        m.induction_matrix(1e-6, pos)
    with pytest.raises(TypeError):
        m.line_induction_matrix(1e-6, pos, cp, None, None)
    with pytest.raises(ValueError):
        m.dual_normal_criterion(1.5, np.zeros((m.n_points, 3)))


def test_dual_normal_criterion_and_free_edges():
    """test/pytests/test_wake_models.py:129-159: one trailing edge found on the wedge."""
    node_positions = np.array([[0, 1, 0], [3, 1, -0.5], [3, 1, +0.5], [0, -1, 0], [3, -1, -0.5], [3, -1, +0.5]],
                              np.float64)
    msh = Mesh(6, [[0, 3, 5, 2], [3, 0, 1, 4], [1, 2, 5, 4]])
    dual = msh.compute_dual()
    t = O.mesh_from_elements(6, [[0, 3, 5, 2], [3, 0, 1, 4], [1, 2, 5, 4]])
    normals = O.surface_normal(t, node_positions)
    te = dual.dual_normal_criterion(-0.5, normals)
    assert len(te) == 1
    ln = msh.get_line(int(te[0]))
    assert (ln.begin, ln.end) in ((0, 3), (3, 0))
    free = dual.dual_free_edges()
    dl = dual.line_data
    assert all(INVALID_ID in (int(dl[i, 0]), int(dl[i, 1])) for i in free)


def test_gpu_entry_fails_loudly_without_device():
    cnt = C.c_int()
    _lib.load().pvl_device_count(C.byref(cnt))
    if cnt.value > 0:
        pytest.skip("a GPU is visible; this checks the no-device path")
    pos, el, _ = G.flat_wing(4, 2, 4.0, 1.0)
    m = Mesh(pos.shape[0], el)
    cp = np.zeros((3, 3))
    with pytest.raises(_lib.PvlError):
        m.induction_matrix3(1e-6, pos, cp, cp)
