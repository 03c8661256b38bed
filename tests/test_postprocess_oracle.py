"""The post-processing oracle on the CPU: Mesh.line_forces against the reference's own C
(src/meshobject.c:1288-1396 compiled unmodified into oracle/_ref/cvl) bit for bit, and
against what the mathematics fixes; the serialisation round trips of the state.
"""

import numpy as np
import pytest

from oracle import oracle as O
from oracle import postprocess_oracle as PO
from oracle import wake_oracle as WO
from pyvl_amd import fio
from pyvl_amd import geometry as G

RNG = np.random.default_rng(21)


def _cases():
    pos, el, _ = G.flat_wing(8, 4, 4.0, 1.0, taper=0.6)
    yield "wing", pos + 0.02 * RNG.normal(size=pos.shape), el
    pos, el = G.uv_sphere(16, 8, sigma=1e-3, seed=3)
    yield "sphere", pos, el
    yield "mixed", np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0.3], [2, 1, 0.1]], float), \
        [[0, 1, 2, 3], [1, 4, 5, 2], [2, 5, 3]]


@pytest.mark.parametrize("name,pos,el", list(_cases()), ids=[c[0] for c in _cases()])
def test_line_forces_oracle_equals_reference(name, pos, el):
    cvl = O.load_ref_cvl()
    if cvl is None:
        pytest.skip("oracle/_ref/cvl (the reference's C extension) is not built")
    topo = O.mesh_from_elements(pos.shape[0], el)
    dual_lines, _, _ = O.mesh_dual(topo)
    circ = RNG.normal(size=topo.n_surfaces)
    vel = RNG.normal(size=pos.shape)
    want = cvl.Mesh.line_forces(cvl.Mesh(pos.shape[0], el), cvl.Mesh(pos.shape[0], el).compute_dual(), circ,
                                np.ascontiguousarray(pos), vel)
    got = O.line_forces(topo, dual_lines, circ, pos, vel)
    assert np.array_equal(got, want)


def test_line_forces_closed_forms():
    """Kutta-Joukowski per line: f = dr x v * c.  A uniform velocity and one panel: the
    forces sum to (sum dr) x v c = 0 round a closed loop; two panels of equal circulation
    cancel on their shared line; the force is linear in the circulation."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0], [2, 1, 0]], float)
    topo = O.mesh_from_elements(6, [[0, 1, 2, 3], [1, 4, 5, 2]])
    dual_lines, _, _ = O.mesh_dual(topo)
    v = np.tile(np.array([1.0, 0.2, -0.3]), (6, 1))
    f1 = O.line_forces(topo, dual_lines, np.array([0.7, 0.0]), pos, v)
    assert np.allclose(f1.sum(axis=0), 0.0, atol=1e-15)
    for il in range(topo.n_lines):
        a, b = topo.lines[il] & O.MASK
        dr = pos[b] - pos[a]
        c = (0.7 if (dual_lines[il, 0] & O.MASK) == 0 else 0.0) - (0.7 if (dual_lines[il, 1] & O.MASK) == 0 else 0.0)
        assert np.allclose(f1[il], np.cross(dr, v[a] * c), atol=1e-15)
    f2 = O.line_forces(topo, dual_lines, np.array([0.5, 0.5]), pos, v)
    shared = [il for il in range(topo.n_lines) if (dual_lines[il] & O.MASK != O.INVALID).all()]
    assert len(shared) == 1 and np.array_equal(f2[shared[0]], np.zeros(3))
    f3 = O.line_forces(topo, dual_lines, np.array([1.4, 0.0]), pos, v)
    assert np.allclose(f3, 2 * f1, rtol=0, atol=1e-15)


def test_correct_forces_index_is_the_reference_line_4i_plus_1():
    """python/pyvl/wake_models.py:317-330 subtracts force 4 i + 1 of the one-row ring mesh:
    the ring [0, 2, 3, 1] + 4 i of _create_wake_mesh(1, n) numbers its lines from the closing
    edge (1, 0) on, so line 4 i + 1 is (0, 2) + 4 i: the row-0 filament from side 0 to side 1,
    the one on the shedding line."""
    tmp = WO.wake_ring_topology(3, 1)
    for i in range(3):
        a, b = tmp.lines[4 * i + 1] & O.MASK
        assert (a, b) == (4 * i, 4 * i + 2)
    # and the oracle's correction touches exactly the shedding lines
    pos, el, _ = G.flat_wing(4, 2, 2.0, 1.0)
    topo = O.mesh_from_elements(pos.shape[0], el)
    shed = np.array([1, 5, 9, 13])
    w = WO.OracleWake(np.array([[0, 1], [1, 2], [2, 3], [3, 4]]), shed, np.array([[0, O.INVALID]] * 4), 1e-6, 0.0, 2)
    w.wake_positions[...] = RNG.normal(size=w.wake_positions.shape)
    w.circulation[...] = RNG.normal(size=w.circulation.shape)
    f = np.zeros((topo.n_lines, 3))
    PO.correct_forces(w, f, topo, pos, RNG.normal(size=topo.n_surfaces), np.array([1.0, 0.0, 0.0]))
    touched = np.nonzero(np.any(f != 0, axis=1))[0]
    assert set(touched.tolist()) <= set(shed.tolist()) and len(touched) > 0


def test_state_maps_round_trip_exactly(tmp_path):
    m = fio.HirearchicalMap()
    m.insert_array("a", RNG.normal(size=(5, 3)))
    m.insert_int("it", 7)
    m.insert_scalar("t", 0.1 + 0.2)
    sub = fio.HirearchicalMap()
    sub.insert_array("w", RNG.normal(size=11))
    sub.insert_string("type", "x.y")
    m.insert_hirearchycal_map("sub", sub)
    with pytest.raises(KeyError):
        m.insert_int("it", 8)
    with pytest.raises(ValueError):
        sub.insert_hirearchycal_map("loop", m)
    for ser, de, name in ((fio.serialize_json, fio.deserialize_json, "s.json"),
                          (fio.serialize_npz, fio.deserialize_npz, "s.npz")):
        ser(m, tmp_path / name)
        r = de(tmp_path / name)
        assert np.array_equal(r.get_array("a"), m.get_array("a"))
        assert r.get_int("it") == 7 and r.get_scalar("t") == 0.1 + 0.2
        assert np.array_equal(r.get_hirearchical_map("sub").get_array("w"), sub.get_array("w"))
        assert r.get_hirearchical_map("sub").get_string("type") == "x.y"
