"""ORACLE -- test infrastructure only.  Restatement of the time step's rigid-frame kinematics.

``kinematics(frame, p0, t)`` follows the chain walk of python/pyvl/solver.py:180-193 (child
frame first: p = R^T (p - d), add_velocity at the new p, v = R^T v) with its own copies of
the reference's arithmetic:
  - rotation from angles: src/core/common.h:165-205 (real3x3_from_angles and its inverse);
  - matrix-vector product: real3x3_vecmul (src/core/common.h:140-147), (m0 x + m1 y) + m2 z;
  - base frame at_time: src/referenceframeobject.c:700-729 (unchanged, parent evolved);
  - TranslatingReferenceFrame.at_time / add_velocity: python/pyvl/reference_frames.py:65-147;
  - RotorReferenceFrame._rotate / at_time / add_velocity: python/pyvl/reference_frames.py:211-322;
  - angle clamp and angles from a matrix: src/core/common.h:207-218.
The frames are read through their public parameters only (offset, angles, parent, vel,
omega, time, and the oscillating frames' amplitudes and phases).  The heaving and pitching
frames, which the reference does not have, are evaluated in closed form,
P = Ry(theta) (p - pivot) + pivot + h z and V = dP/dt, as the innermost frame of a chain.

Pinned by tests/test_frames.py: transforms and at_time equal to the reference's own
ReferenceFrame (oracle/_ref/cvl, its C compiled unmodified) bit for bit; the rotor's at_time
against Rodrigues' rotation; translating offsets in closed form; a planted mistake rejected.
"""

from __future__ import annotations

import math

import numpy as np


def clamp(a: float) -> float:
    rem = math.remainder(float(a), 2 * math.pi)
    return 2 * math.pi + rem if rem < 0 else rem


def mat_from_angles(a) -> np.ndarray:
    cx, sx = math.cos(a[0]), math.sin(a[0])
    cy, sy = math.cos(a[1]), math.sin(a[1])
    cz, sz = math.cos(a[2]), math.sin(a[2])
    return np.array([[cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz],
                     [cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz],
                     [-sy, sx * cy, cx * cy]], np.float64)


def vecmul(m, p) -> np.ndarray:
    out = np.empty_like(p)
    for i in range(3):
        out[..., i] = (m[i, 0] * p[..., 0] + m[i, 1] * p[..., 1]) + m[i, 2] * p[..., 2]
    return out


def angles_from(m) -> np.ndarray:
    return np.array([math.atan2(m[2, 1], m[2, 2]), math.atan2(-m[2, 0], math.hypot(m[0, 0], m[1, 0])),
                     math.atan2(m[1, 0], m[0, 0])], np.float64)


def rotor_rotate(omega, r, dt):
    """python/pyvl/reference_frames.py:211-243 as written."""
    angular_velocity = np.linalg.norm(omega)
    if angular_velocity == 0:
        return np.array(r)
    direction = omega / angular_velocity
    rp = np.dot(r, direction) * direction
    v1 = r - rp
    v2 = np.cross(r, direction)
    phi = np.arctan2(v2, v1)
    rad = np.hypot(v1, v2)
    return rp + rad * np.cos(phi + angular_velocity * dt)


class _Evolved:
    """One frame of the chain at time t: (angles, offset, kind, term)."""

    def __init__(self, angles, offset, kind, term, parent, reclamp=True):
        # a new frame object clamps its angles (its constructor); the base frame's at_time
        # copies the transformation as it is
        self.angles = np.array([clamp(x) if reclamp else float(x) for x in angles], np.float64)
        self.offset = np.asarray(offset, np.float64)
        self.kind = kind
        self.term = term
        self.parent = parent


def evolve(frame, t: float) -> _Evolved | None:
    """The chain's at_time, from the frames' public parameters."""
    if frame is None:
        return None
    name = type(frame).__name__
    parent = evolve(frame.parent, t) if frame.parent is not None else None
    if name == "TranslatingReferenceFrame":
        dt = t - frame.time
        return _Evolved(frame.angles, frame.offset + dt * np.asarray(frame.vel, np.float64), "translating",
                        np.asarray(frame.vel, np.float64), parent)
    if name == "RotorReferenceFrame":
        rot = mat_from_angles(frame.angles)
        omega = np.asarray(frame.omega, np.float64)
        for i in range(3):
            rot[i, :] = rotor_rotate(omega, rot[i, :], t - frame.time)
        return _Evolved(angles_from(rot), frame.offset, "rotor", omega, parent)
    if name in ("HeaveFrame", "PitchHeaveFrame"):
        if frame.parent is not None:
            raise NotImplementedError("closed-form frames are restated as the innermost frame of a chain only")
        return _Evolved((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), name, frame, None)
    return _Evolved(frame.angles, frame.offset, "none", None, parent, reclamp=False)


def _closed_form(kind: str, f, p0, t):
    if kind == "HeaveFrame":
        pos = np.array(p0, np.float64)
        pos[:, 2] += f.h0 * np.sin(f.omega * t + f.phase)
        vel = np.zeros_like(pos)
        vel[:, 2] = f.h0 * f.omega * np.cos(f.omega * t + f.phase)
        return pos, vel
    th = f.theta0 * np.sin(f.omega * t + f.phase_p)
    thd = f.theta0 * f.omega * np.cos(f.omega * t + f.phase_p)
    h = f.h0 * np.sin(f.omega * t + f.phase_h)
    hd = f.h0 * f.omega * np.cos(f.omega * t + f.phase_h)
    c, s = np.cos(th), np.sin(th)
    x = p0[:, 0] - f.pivot_x
    z = p0[:, 2]
    pos = np.array(p0, np.float64)
    pos[:, 0] = f.pivot_x + c * x + s * z
    pos[:, 2] = -s * x + c * z + h
    vel = np.zeros_like(pos)
    vel[:, 0] = thd * (-s * x + c * z)
    vel[:, 2] = thd * (-c * x - s * z) + hd
    return pos, vel


def kinematics(frame, p0, t: float) -> tuple[np.ndarray, np.ndarray]:
    """Positions and velocities of a body's points in the global frame at time t.  A frame
    object with its own closed form (``positions(p0, t)`` / ``velocities(p0, t)``, as test
    fixtures define) is evaluated through it."""
    if hasattr(frame, "positions") and hasattr(frame, "velocities"):
        return frame.positions(p0, t), frame.velocities(p0, t)
    f = evolve(frame, float(t))
    pos = np.array(p0, np.float64)
    vel = np.zeros_like(pos)
    while f is not None:
        if f.kind in ("HeaveFrame", "PitchHeaveFrame"):
            pos, vel = _closed_form(f.kind, f.term, pos, float(t))
            f = f.parent
            continue
        rinv = np.ascontiguousarray(mat_from_angles(f.angles).T)
        pos = vecmul(rinv, pos - f.offset)
        if f.kind == "translating":
            vel = vel + f.term[None, :]
        elif f.kind == "rotor":
            vel = vel + np.cross(pos, f.term[None, :])
        vel = vecmul(rinv, vel)
        f = f.parent
    return pos, vel


def _oscillating_add_velocity(f, t, pos, vel):
    """add_velocity of the heaving / pitching frames at their evaluated time: the heave rate
    along z, or R (dP/dt) written as W p + c (the form the time step's chain walk turns into
    dP/dt, pinned through the closed forms above)."""
    if type(f).__name__ == "HeaveFrame":
        vel[:, 2] += f.h0 * f.omega * math.cos(f.omega * t + f.phase)
        return
    th = f.theta0 * math.sin(f.omega * t + f.phase_p)
    thd = f.theta0 * f.omega * math.cos(f.omega * t + f.phase_p)
    h = f.h0 * math.sin(f.omega * t + f.phase_h)
    hd = f.h0 * f.omega * math.cos(f.omega * t + f.phase_h)
    c, s = math.cos(th), math.sin(th)
    j = np.array([[0.0, 0.0, thd], [0.0, 0.0, 0.0], [-thd, 0.0, 0.0]])
    r = np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])
    w = r @ j
    cv = r @ (np.array([0.0, 0.0, hd]) - j @ np.array([f.pivot_x, 0.0, h]))
    vel[:, :] += vecmul(w, pos) + cv[None, :]


def f_h(frame, t):
    return frame.h0 * math.sin(frame.omega * t + frame.phase)


def _frame_state(frame, t):
    """(angles, offset) of a frame after at_time(t), and its velocity rule."""
    name = type(frame).__name__
    if name == "HeaveFrame":
        # the frame displaced by -h(t) along z, unrotated
        return np.zeros(3), np.array([0.0, 0.0, -f_h(frame, t)]), ("osc", frame)
    if name == "PitchHeaveFrame":
        # angles (0, -theta, 0); offset d with R^T (p - d) = Ry(theta) (p - pivot) + pivot + h z
        th = frame.theta0 * math.sin(frame.omega * t + frame.phase_p)
        h = frame.h0 * math.sin(frame.omega * t + frame.phase_h)
        c, s = math.cos(th), math.sin(th)
        qx, qz = frame.pivot_x - c * frame.pivot_x, s * frame.pivot_x + h
        return (np.array([0.0, clamp(-th), 0.0]), np.array([-(c * qx - s * qz), 0.0, -(s * qx + c * qz)]),
                ("osc", frame))
    ev = evolve(frame, t)
    return ev.angles, ev.offset, (ev.kind, ev.term)


def positions_at_time(frame, p0, t):
    """SimulationGeometry.positions_at_time (python/pyvl/geometry.py:518-539): the evolved
    frame's to_global_with_offset (src/referenceframeobject.c:556-582, with its merge of
    src/core/transformation.h:57-62 starting from the frame itself)."""
    chain = []
    f = frame
    while f is not None:
        chain.append(_frame_state(f, t))
        f = f.parent
    mat = np.ascontiguousarray(mat_from_angles(chain[0][0]).T)
    off = np.array(chain[0][1], np.float64)
    for k in range(len(chain) - 1):
        ta = np.ascontiguousarray(mat_from_angles(chain[k][0]).T)
        off_a = np.array(chain[k][1], np.float64)
        new_off = vecmul(ta, off_a + vecmul(mat, off))
        new_mat = np.empty((3, 3))
        for i in range(3):
            for j in range(3):
                new_mat[i, j] = (mat[i, 0] * ta[0, j] + mat[i, 1] * ta[1, j]) + mat[i, 2] * ta[2, j]
        mat, off = new_mat, new_off
    return vecmul(mat, np.asarray(p0, np.float64)) - off


def velocity_at_time(frame, p0, t):
    """SimulationGeometry.velocity_at_time (python/pyvl/geometry.py:541-570): add_velocity at
    the current positions, then both to the parent."""
    pos = np.array(p0, np.float64)
    vel = np.zeros_like(pos)
    f = frame
    while f is not None:
        angles, offset, (kind, term) = _frame_state(f, t)
        if kind == "translating":
            vel = vel + term[None, :]
        elif kind == "rotor":
            vel = vel + np.cross(pos, term[None, :])
        elif kind == "osc":
            _oscillating_add_velocity(term, t, pos, vel)
        rinv = np.ascontiguousarray(mat_from_angles(angles).T)
        pos = vecmul(rinv, pos - offset)
        vel = vecmul(rinv, vel)
        f = f.parent
    return vel
