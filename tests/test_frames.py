"""Reference frames on the CPU: the Python mirror and the oracle's chain walk against the
reference's own ReferenceFrame (src/referenceframeobject.c compiled unmodified into
oracle/_ref/cvl), and against what the mathematics fixes for the frames the reference
writes in Python (translating and rotor frames, python/pyvl/reference_frames.py) and for
the oscillating frames it does not have.
"""

import math

import numpy as np
import pytest

from oracle import frames_oracle as FO
from oracle import oracle as O
from pyvl_amd import reference_frames as RF

RNG = np.random.default_rng(11)


def _cvl():
    m = O.load_ref_cvl()
    if m is None:
        pytest.skip("oracle/_ref/cvl (the reference's C extension) is not built")
    return m


def _params(k):
    rng = np.random.default_rng(100 + k)
    return tuple(rng.normal(0, 2, 3)), tuple(rng.uniform(-7, 7, 3))


def _chains(cvl, depth):
    """The same chain built as the reference's frames and as this package's."""
    ref, own = None, None
    for k in range(depth):
        off, ang = _params(k)
        ref = cvl.ReferenceFrame(off, ang, ref)
        own = RF.ReferenceFrame(off, ang, own)
    return ref, own


@pytest.mark.parametrize("depth", [1, 2, 4])
def test_transforms_equal_reference_frame(depth):
    cvl = _cvl()
    ref, own = _chains(cvl, depth)
    pts = RNG.normal(0, 3, (257, 3))
    assert np.array_equal(ref.angles, own.angles)
    assert np.array_equal(ref.offset, own.offset)
    assert np.array_equal(ref.rotation_matrix, own.rotation_matrix)
    assert np.array_equal(ref.rotation_matrix_inverse, own.rotation_matrix_inverse)
    for name in ("from_parent_with_offset", "from_parent_without_offset", "to_parent_with_offset",
                 "to_parent_without_offset", "from_global_with_offset", "from_global_without_offset",
                 "to_global_with_offset", "to_global_without_offset"):
        want = getattr(ref, name)(pts)
        got = getattr(own, name)(pts)
        assert np.array_equal(got, want), name
        out = np.empty_like(pts)
        assert getattr(own, name)(pts, out) is out and np.array_equal(out, want), name
    for axis in "xyz":
        r2, o2 = getattr(ref, "rotate_" + axis)(0.7), getattr(own, "rotate_" + axis)(0.7)
        assert np.array_equal(r2.angles, o2.angles)
    assert np.array_equal(ref.with_offset((1.0, -2.0, 0.5)).offset, own.with_offset((1.0, -2.0, 0.5)).offset)
    m = own.rotation_matrix
    assert np.array_equal(cvl.ReferenceFrame.angles_from_rotation(m), RF.ReferenceFrame.angles_from_rotation(m))
    rt, ot = ref.at_time(3.25), own.at_time(3.25)
    assert type(ot) is RF.ReferenceFrame and np.array_equal(rt.to_global_with_offset(pts), ot.to_global_with_offset(pts))
    assert len(ref.parents) == len(own.parents) == depth - 1
    assert repr(ref).replace("pyvl.cvl.", "") == repr(own)


def test_transform_argument_errors_match_reference():
    cvl = _cvl()
    ref, own = cvl.ReferenceFrame(), RF.ReferenceFrame()
    for args, exc in (((np.zeros((3, 2)),), ValueError), ((np.zeros((4, 3)), [1, 2]), TypeError),
                      ((np.zeros((4, 3)), np.zeros((3, 3))), ValueError),
                      ((np.zeros((4, 3)), np.zeros((4, 3), np.float32)), ValueError)):
        with pytest.raises(exc):
            ref.to_parent_with_offset(*args)
        with pytest.raises(exc):
            own.to_parent_with_offset(*args)


def _reference_chain_walk(frame, p0, t):
    """python/pyvl/solver.py:180-193 on the reference's own frames."""
    rf = frame.at_time(t)
    pos = np.array(p0)
    vel = np.zeros_like(pos)
    while rf is not None:
        pos = rf.to_parent_with_offset(pos, pos)
        rf.add_velocity(pos, vel)
        vel = rf.to_parent_without_offset(vel, vel)
        rf = rf.parent
    return pos, vel


def test_oracle_chain_walk_equals_reference_on_base_frames():
    cvl = _cvl()
    ref, own = _chains(cvl, 3)
    p0 = RNG.normal(size=(100, 3))
    pr, vr = _reference_chain_walk(ref, p0, 0.4)
    po, vo = FO.kinematics(own, p0, 0.4)
    assert np.array_equal(po, pr) and np.array_equal(vo, vr)


def _rodrigues(axis, angle):
    k = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + math.sin(angle) * kx + (1 - math.cos(angle)) * (kx @ kx)


def test_rotor_at_time_is_a_rotation_about_omega():
    omega = np.array([0.3, -1.1, 0.7])
    f = RF.RotorReferenceFrame((1.0, 2.0, 3.0), (0.2, 0.4, -0.3), None, omega, 0.5)
    for t in (0.5, 0.9, 3.7):
        want = f.rotation_matrix @ _rodrigues(omega, np.linalg.norm(omega) * (t - 0.5)).T
        got_own = f.at_time(t).rotation_matrix
        got_orc = FO.mat_from_angles(FO.evolve(f, t).angles)
        assert np.max(np.abs(got_own - want)) < 1e-13
        assert np.max(np.abs(got_orc - want)) < 1e-13
        assert np.array_equal(f.at_time(t).offset, f.offset)


def test_translating_at_time_closed_form():
    f = RF.TranslatingReferenceFrame((1.0, -2.0, 0.5), (0.0, 0.3, 0.0), (2.0, 0.5, -1.0), None, 1.0)
    for t in (1.0, 1.25, -3.0):
        assert np.array_equal(f.at_time(t).offset, np.array([1.0, -2.0, 0.5]) + (t - 1.0) * np.array([2.0, 0.5, -1.0]))
        assert np.array_equal(FO.evolve(f, t).offset, f.at_time(t).offset)
    p0 = RNG.normal(size=(20, 3))
    pos, vel = FO.kinematics(f, p0, 2.0)
    rinv = f.rotation_matrix_inverse
    assert np.allclose(vel, np.tile(rinv @ np.array([2.0, 0.5, -1.0]), (20, 1)), rtol=0, atol=1e-15)
    # positions: R^T (p - d(t)), d(t) = d + (t - t0) v, written out
    d = np.array([1.0, -2.0, 0.5]) + 1.0 * np.array([2.0, 0.5, -1.0])
    assert np.allclose(pos, (p0 - d) @ rinv.T, rtol=0, atol=1e-14)
    # (to_global_with_offset is R^T p - d for a single frame, src/referenceframeobject.c:570-573:
    # not the time step's chain walk once the frame is rotated)
    assert not np.allclose(pos, f.at_time(2.0).to_global_with_offset(p0), rtol=0, atol=1e-3)


def _mixed_chain():
    root = RF.RotorReferenceFrame((0.5, 0.0, -1.0), (0.1, 0.0, 0.2), None, (0.0, 0.2, 1.5), 0.0)
    mid = RF.TranslatingReferenceFrame((0.0, 1.0, 0.0), (0.0, 0.5, 0.0), (1.0, 0.0, 0.25), root, 0.25)
    return RF.ReferenceFrame((0.3, -0.2, 0.1), (1.0, 2.0, 3.0), mid)


@pytest.mark.parametrize("t", [0.0, 0.3, 2.5])
def test_host_chain_walk_equals_oracle(t):
    p0 = RNG.normal(size=(64, 3))
    f = _mixed_chain()
    pos, vel = RF.frame_chain_kinematics(f, p0, t)
    po, vo = FO.kinematics(f, p0, t)
    assert np.array_equal(pos, po) and np.array_equal(vel, vo)


def test_rotor_velocity_is_omega_cross():
    """A rotor frame alone: positions rotated about omega, velocity (R^T of) p x omega at
    the rotated positions, the reference's convention (python/pyvl/reference_frames.py:298-322)."""
    omega = np.array([0.0, 0.0, 2.0])
    f = RF.RotorReferenceFrame(omega=omega)
    p0 = RNG.normal(size=(10, 3))
    t = 0.3
    pos, vel = FO.kinematics(f, p0, t)
    rot = _rodrigues(omega, 2.0 * t)
    assert np.allclose(pos, p0 @ rot.T, rtol=0, atol=1e-14)  # R^T = Rodrigues: rows rotated by +w t
    want = (np.cross(pos, omega)) @ rot.T
    assert np.allclose(vel, want, rtol=0, atol=1e-14)


@pytest.mark.parametrize("kind", ["heave", "pitch_heave"])
def test_oscillating_frames_match_their_closed_forms(kind):
    if kind == "heave":
        f = RF.HeaveFrame(0.1, 0.6, 0.2)
    else:
        f = RF.PitchHeaveFrame(np.radians(8.0), 0.12, 0.8, 0.25, 0.3, 1.1)
    p0 = RNG.normal(size=(40, 3))
    for t in (0.0, 0.77, 5.0):
        pos, vel = RF.frame_chain_kinematics(f, p0, t)
        po, vo = FO.kinematics(f, p0, t)
        assert np.max(np.abs(pos - po)) <= 4e-15 * (1 + np.max(np.abs(po)))
        assert np.max(np.abs(vel - vo)) <= 4e-15 * (1 + np.max(np.abs(vo)))
        # velocity is the time derivative of position (central difference)
        h = 1e-5
        fd = (FO.kinematics(f, p0, t + h)[0] - FO.kinematics(f, p0, t - h)[0]) / (2 * h)
        assert np.max(np.abs(fd - vo)) < 1e-8


def test_planted_rotation_mistake_is_caught():
    """A chain walk that applies R instead of R^T disagrees with the reference's frames."""
    cvl = _cvl()
    ref, own = _chains(cvl, 2)
    p0 = RNG.normal(size=(30, 3))
    pr, _ = _reference_chain_walk(ref, p0, 0.0)
    f, pos = own, p0.copy()
    while f is not None:
        pos = FO.vecmul(FO.mat_from_angles(f.angles), pos - f.offset)
        f = f.parent
    assert np.max(np.abs(pos - pr)) > 1e-3


def test_device_terms_required_for_custom_velocity():
    class Custom(RF.ReferenceFrame):
        def add_velocity(self, positions, velocity, /):
            velocity += 1.0

    with pytest.raises(TypeError):
        RF.device_chain_terms(Custom())
    terms, kinds = RF.device_chain_terms(_mixed_chain().at_time(0.5))
    assert terms.shape == (3, 24) and list(kinds) == [RF.KIND_NONE, RF.KIND_TRANSLATING, RF.KIND_ROTOR]
