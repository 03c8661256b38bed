"""Reference frames: the kinematics plug-in contract of the time step.

Mirrors the reference's ``pyvl.cvl.ReferenceFrame`` (src/referenceframeobject.c; rotation
from angles src/core/common.h:165-214, merged transformations src/core/transformation.h)
and its Python subclasses ``TranslatingReferenceFrame`` / ``RotorReferenceFrame``
(python/pyvl/reference_frames.py:15-322): the same constructor arguments, properties,
transformations, ``at_time`` and ``add_velocity`` contract.  Two oscillating frames for the
heaving and pitching wing workloads (which the reference does not have) plug into the same
contract: :class:`HeaveFrame` and :class:`PitchHeaveFrame`.

Every frame also describes itself to the device (:meth:`ReferenceFrame.device_terms`): the
chain walk of the time step (python/pyvl/solver.py:180-193) runs on the GPU as one kernel
over all points (``pvl_frame_kinematics``), with the frames' rotation matrices, offsets and
velocity terms evaluated once per step on the host.

Transformations evaluate each matrix-vector product as the reference's real3x3_vecmul,
(m0 x + m1 y) + m2 z, without fused multiply-adds.
"""

from __future__ import annotations

import math

import numpy as np

# Velocity-term kinds understood by the device kinematics kernel.
KIND_NONE = 0         # the base frame adds nothing
KIND_TRANSLATING = 1  # v += c
KIND_ROTOR = 2        # v += p x w
KIND_AFFINE = 3       # v += W p + c (oscillating frames)


def clamp_angle_to_range(a: float) -> float:
    """src/core/common.h:212-218: IEEE remainder modulo 2 pi, shifted into [0, 2 pi)."""
    rem = math.remainder(float(a), 2 * math.pi)
    if rem < 0:
        return 2 * math.pi + rem
    return rem


def matrix_from_angles(angles) -> np.ndarray:
    """real3x3_from_angles (src/core/common.h:165-185)."""
    ax, ay, az = (float(v) for v in angles)
    cx, sx = math.cos(ax), math.sin(ax)
    cy, sy = math.cos(ay), math.sin(ay)
    cz, sz = math.cos(az), math.sin(az)
    return np.array([
        [cy * cz, sx * sy * cz - cx * sz, cx * sy * cz + sx * sz],
        [cy * sz, sx * sy * sz + cx * cz, cx * sy * sz - sx * cz],
        [-sy, sx * cy, cx * cy],
    ], np.float64)


def inverse_matrix_from_angles(angles) -> np.ndarray:
    """real3x3_inverse_from_angles (src/core/common.h:186-205): the transpose, with each
    entry computed by the same expression."""
    return np.ascontiguousarray(matrix_from_angles(angles).T)


def angles_from_matrix(m) -> np.ndarray:
    """angles_from_real3x3 (src/core/common.h:207-210)."""
    m = np.asarray(m, np.float64).reshape(3, 3)
    return np.array([math.atan2(m[2, 1], m[2, 2]), math.atan2(-m[2, 0], math.hypot(m[0, 0], m[1, 0])),
                     math.atan2(m[1, 0], m[0, 0])], np.float64)


def _matmul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """real3x3_matmul (src/core/common.h:148-164): each entry a row . column dot product."""
    out = np.empty((3, 3), np.float64)
    for i in range(3):
        for j in range(3):
            out[i, j] = (a[i, 0] * b[0, j] + a[i, 1] * b[1, j]) + a[i, 2] * b[2, j]
    return out


def _vecmul_rows(m: np.ndarray, p: np.ndarray) -> np.ndarray:
    """real3x3_vecmul over rows of p (..., 3)."""
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    out = np.empty_like(p)
    for i in range(3):
        out[..., i] = (m[i, 0] * x + m[i, 1] * y) + m[i, 2] * z
    return out


def _vec3(v, what: str) -> tuple[float, float, float]:
    try:
        t = tuple(float(x) for x in v)
    except TypeError as e:
        raise TypeError(f"{what} must be a sequence of 3 numbers") from e
    if len(t) != 3:
        raise TypeError(f"{what} must be a sequence of 3 numbers")
    return t


def _prepare(args, nargs_ok=(1, 2)):
    """prepare_for_transformation (src/referenceframeobject.c:284-366)."""
    if len(args) not in nargs_ok:
        raise TypeError("Method only takes 1 or two parameters")
    if len(args) == 2 and not isinstance(args[1], np.ndarray):
        raise TypeError("Second argument must be an array.")
    inp = np.ascontiguousarray(np.asarray(args[0], dtype=np.float64))
    if inp.ndim < 1:
        raise ValueError("Input array must have at least one axis")
    if inp.shape[-1] != 3:
        raise ValueError(
            "Input array does not have the last axis with 3 dimensions "
            f"(shape is (..., {inp.shape[-1]}) instead of (..., 3)).")
    if len(args) == 2:
        out = args[1]
        if not out.flags.c_contiguous:
            raise ValueError("Output array is not C-contiguous.")
        if out.dtype != np.float64:
            raise ValueError("Output array is not of C double type.")
        if out.ndim != inp.ndim:
            raise ValueError("Output array does not have the same number of axis as the input "
                             f"array (in: {inp.ndim}, out: {out.ndim}).")
        if out.shape != inp.shape:
            raise ValueError("Output array does not have the same exact shape as the input array.")
    else:
        out = np.empty_like(inp)
    return inp, out


class ReferenceFrame:
    """A rotation (angles about x, y, z) and an offset relative to an optional parent frame.

    ``from_parent``: x -> R x + d; ``to_parent``: x -> R^T (x - d)
    (src/referenceframeobject.c:367-452).
    """

    def __init__(self, offset=(0.0, 0.0, 0.0), theta=(0.0, 0.0, 0.0), parent: "ReferenceFrame | None" = None):
        if parent is not None and not isinstance(parent, ReferenceFrame):
            raise TypeError(f'Argument "parent" must be a ReferenceFrame object, but it was {type(parent)!r}')
        ox, oy, oz = _vec3(offset, "offset")
        tx, ty, tz = _vec3(theta, "theta")
        self._offset = np.array([ox, oy, oz], np.float64)
        self._angles = np.array([clamp_angle_to_range(tx), clamp_angle_to_range(ty), clamp_angle_to_range(tz)],
                                np.float64)
        self._parent = parent

    # -- properties (src/referenceframeobject.c:95-279) --------------------------------
    @property
    def parent(self) -> "ReferenceFrame | None":
        return self._parent

    @property
    def offset(self) -> np.ndarray:
        return self._offset.copy()

    @property
    def angles(self) -> np.ndarray:
        return self._angles.copy()

    @property
    def rotation_matrix(self) -> np.ndarray:
        return matrix_from_angles(self._angles)

    @property
    def rotation_matrix_inverse(self) -> np.ndarray:
        return inverse_matrix_from_angles(self._angles)

    @property
    def parents(self) -> tuple:
        out = []
        p = self._parent
        while p is not None:
            out.append(p)
            p = p._parent
        return tuple(out)

    def __repr__(self) -> str:
        a, o = self._angles, self._offset
        body = f"({a[0]:g}, {a[1]:g}, {a[2]:g}), ({o[0]:g}, {o[1]:g}, {o[2]:g})"
        if self._parent is not None:
            return f"ReferenceFrame({body}, parent={self._parent!r})"
        return f"ReferenceFrame({body})"

    # -- transformations (src/referenceframeobject.c:367-610) ---------------------------
    def from_parent_with_offset(self, *args) -> np.ndarray:
        inp, out = _prepare(args)
        out[...] = _vecmul_rows(self.rotation_matrix, inp) + self._offset
        return out

    def from_parent_without_offset(self, *args) -> np.ndarray:
        inp, out = _prepare(args)
        out[...] = _vecmul_rows(self.rotation_matrix, inp)
        return out

    def to_parent_with_offset(self, *args) -> np.ndarray:
        inp, out = _prepare(args)
        out[...] = _vecmul_rows(self.rotation_matrix_inverse, inp - self._offset)
        return out

    def to_parent_without_offset(self, *args) -> np.ndarray:
        inp, out = _prepare(args)
        out[...] = _vecmul_rows(self.rotation_matrix_inverse, inp)
        return out

    def from_global_with_offset(self, *args) -> np.ndarray:
        inp, out = _prepare(args)
        mat, off = self.rotation_matrix, self._offset.copy()
        p = self
        while p._parent is not None:
            # merge_transformations(T_p, r_p, mat, off) (src/core/transformation.h:33-38)
            tp = p.rotation_matrix
            off = p._offset + _vecmul_rows(tp, off)
            mat = _matmul(tp, mat)
            p = p._parent
        out[...] = _vecmul_rows(mat, inp) + off
        return out

    def from_global_without_offset(self, *args) -> np.ndarray:
        inp, out = _prepare(args)
        mat = self.rotation_matrix
        p = self
        while p._parent is not None:
            mat = _matmul(p.rotation_matrix, mat)
            p = p._parent
        out[...] = _vecmul_rows(mat, inp)
        return out

    def to_global_with_offset(self, *args) -> np.ndarray:
        inp, out = _prepare(args)
        mat, off = self.rotation_matrix_inverse, self._offset.copy()
        p = self
        while p._parent is not None:
            # merge_transformations_reverse(T_p^T, r_p, mat, off) (src/core/transformation.h:57-62)
            ta = p.rotation_matrix_inverse
            off = _vecmul_rows(ta, p._offset + _vecmul_rows(mat, off))
            mat = _matmul(mat, ta)
            p = p._parent
        out[...] = _vecmul_rows(mat, inp) - off
        return out

    def to_global_without_offset(self, *args) -> np.ndarray:
        inp, out = _prepare(args)
        mat = self.rotation_matrix_inverse
        p = self
        while p._parent is not None:
            mat = _matmul(mat, p.rotation_matrix_inverse)
            p = p._parent
        out[...] = _vecmul_rows(mat, inp)
        return out

    # -- derived frames (src/referenceframeobject.c:612-700) -----------------------------
    def _copy_base(self) -> "ReferenceFrame":
        new = ReferenceFrame.__new__(ReferenceFrame)
        new._offset = self._offset.copy()
        new._angles = self._angles.copy()
        new._parent = self._parent
        return new

    def _rotated(self, axis: int, theta: float) -> "ReferenceFrame":
        new = self._copy_base()
        new._angles[axis] = clamp_angle_to_range(new._angles[axis] + float(theta))
        return new

    def rotate_x(self, theta: float) -> "ReferenceFrame":
        return self._rotated(0, theta)

    def rotate_y(self, theta: float) -> "ReferenceFrame":
        return self._rotated(1, theta)

    def rotate_z(self, theta: float) -> "ReferenceFrame":
        return self._rotated(2, theta)

    def with_offset(self, offset) -> "ReferenceFrame":
        off = np.ascontiguousarray(np.asarray(offset, np.float64))
        if off.ndim != 1 or off.size != 3:
            raise ValueError(f"Input array must have 3 element, instead {off.size} were given.")
        new = self._copy_base()
        new._offset = off.copy()
        return new

    def at_time(self, t: float) -> "ReferenceFrame":
        """The base frame does not move; its parent is evolved
        (src/referenceframeobject.c:700-729).  Returns a base ReferenceFrame."""
        t = float(t)
        new = self._copy_base()
        new._parent = self._parent.at_time(t) if self._parent is not None else None
        return new

    @staticmethod
    def angles_from_rotation(mat) -> np.ndarray:
        m = np.ascontiguousarray(np.asarray(mat, np.float64))
        if m.shape != (3, 3):
            raise ValueError(f"Array was not a (3, 3) array, instead it was {m.shape}.")
        return angles_from_matrix(m)

    def add_velocity(self, positions, velocity, /) -> None:
        """The base frame adds nothing; the arrays are only checked
        (src/referenceframeobject.c:939-955)."""
        p = np.asarray(positions)
        if not isinstance(positions, np.ndarray) or p.ndim != 2 or p.shape[1] != 3 or p.dtype != np.float64:
            raise ValueError("Position array must be a (N, 3) float64 array")
        if (not isinstance(velocity, np.ndarray) or velocity.shape != p.shape or velocity.dtype != np.float64
                or not velocity.flags.writeable):
            raise ValueError("Velocity array must be a writable float64 array shaped like the positions")

    # -- serialisation (src/referenceframeobject.c:776-937) ----------------------------
    def save(self, group) -> None:
        group["offset"] = self.offset
        group["angles"] = self.angles
        group["type"] = "pyvl.cvl.ReferenceFrame"

    @classmethod
    def load(cls, group, parent: "ReferenceFrame | None" = None) -> "ReferenceFrame":
        if parent is not None and not isinstance(parent, ReferenceFrame):
            raise TypeError(f"Parent was neither None nor a ReferenceFrame (instead it was {type(parent)!r}).")
        off = np.asarray(group["offset"], np.float64).reshape(-1)
        ang = np.asarray(group["angles"], np.float64).reshape(-1)
        if off.size != 3:
            raise ValueError(f"Offset array did not have 3 elements, but had {off.size} instead.")
        if ang.size != 3:
            raise ValueError(f"Angle array did not have 3 elements, but had {ang.size} instead.")
        new = ReferenceFrame.__new__(cls)
        new._offset = off.copy()
        new._angles = ang.copy()
        new._parent = parent
        return new

    # -- device description ------------------------------------------------------------
    def device_terms(self) -> tuple[int, np.ndarray, np.ndarray]:
        """(kind, W (3,3), c (3,)) of this frame's add_velocity: v += 0 / c / p x w / W p + c,
        evaluated at the parent-frame positions as the time step does."""
        return KIND_NONE, np.zeros((3, 3)), np.zeros(3)


class TranslatingReferenceFrame(ReferenceFrame):
    """Constant linear velocity: r'(t) = R r + d + v (t - t0)
    (python/pyvl/reference_frames.py:15-147)."""

    def __init__(self, offset=(0, 0, 0), theta=(0, 0, 0), velocity=(0, 0, 0), parent=None, time: float = 0.0):
        super().__init__(offset=offset, theta=theta, parent=parent)
        self.vel = np.array(velocity, np.float64).reshape((3,))
        self.time = float(time)

    def at_time(self, t: float) -> "TranslatingReferenceFrame":
        dt = t - self.time
        return TranslatingReferenceFrame(
            offset=self.offset + dt * self.vel, velocity=self.vel, time=t,
            parent=self.parent.at_time(t) if self.parent is not None else None, theta=self.angles)

    def save(self, group) -> None:
        group["velocity"] = self.vel.copy()
        group["time"] = self.time
        super().save(group)

    @classmethod
    def load(cls, group, parent=None) -> "TranslatingReferenceFrame":
        return cls(np.asarray(group["offset"]), np.asarray(group["angles"]), np.asarray(group["velocity"]), parent,
                   float(group["time"]))

    def add_velocity(self, positions, velocity, /) -> None:
        super().add_velocity(positions, velocity)
        velocity[:, :] += self.vel[None, :]

    def device_terms(self):
        return KIND_TRANSLATING, np.zeros((3, 3)), self.vel.copy()


class RotorReferenceFrame(ReferenceFrame):
    """Constant angular velocity omega about the frame's origin; dr/dt = r x omega
    (python/pyvl/reference_frames.py:150-322)."""

    def __init__(self, offset=(0, 0, 0), theta=(0, 0, 0), parent=None, omega=(0, 0, 0), time: float = 0.0):
        super().__init__(offset=offset, theta=theta, parent=parent)
        self.omega = np.array(omega, np.float64).reshape((3,))
        self.time = time

    @staticmethod
    def _rotate(omega, r, dt: float) -> np.ndarray:
        """python/pyvl/reference_frames.py:211-243, element-wise as written there."""
        angular_velocity = np.linalg.norm(omega)
        if angular_velocity == 0:
            return np.array(r)
        direction = omega / angular_velocity
        rp = np.dot(r, direction) * direction
        v1 = r - rp
        v2 = np.cross(r, direction)
        phi = np.arctan2(v2, v1)
        rad = np.hypot(v1, v2)
        dr = rad * np.cos(phi + angular_velocity * dt)
        return rp + dr

    def at_time(self, t: float) -> "RotorReferenceFrame":
        rot_mat = self.rotation_matrix
        for i in range(3):
            rot_mat[i, :] = RotorReferenceFrame._rotate(self.omega, rot_mat[i, :], t - self.time)
        new_angles = ReferenceFrame.angles_from_rotation(rot_mat)
        return RotorReferenceFrame(self.offset, new_angles, self.parent.at_time(t) if self.parent else None,
                                   self.omega, t)

    def __repr__(self) -> str:
        o, a = self.offset, self.angles
        return (f"{type(self).__name__}({o[0]:g}, {o[1]:g}, {o[2]:g}, {a[0]:g}, {a[1]:g}, {a[2]:g}, {self.parent}, "
                f"time={self.time:g}, omega={self.omega})")

    def save(self, group) -> None:
        group["offset"] = self.offset
        group["angles"] = self.angles
        group["omega"] = self.omega.copy()
        group["time"] = self.time

    @classmethod
    def load(cls, group, parent=None) -> "RotorReferenceFrame":
        return cls(offset=np.asarray(group["offset"]), theta=np.asarray(group["angles"]), parent=parent,
                   omega=np.asarray(group["omega"]), time=float(group["time"]))

    def add_velocity(self, positions, velocity, /) -> None:
        super().add_velocity(positions, velocity)
        velocity[:, :] += np.cross(positions, self.omega[None, :])

    def device_terms(self):
        return KIND_ROTOR, np.zeros((3, 3)), self.omega.copy()


class HeaveFrame(ReferenceFrame):
    """Vertical oscillation z(t) = h0 sin(omega t + phase) of the body (no reference
    counterpart; the heaving-wing workload).  ``at_time`` gives the frame displaced by
    -h(t) along z (to_parent subtracts the offset), ``add_velocity`` the heave rate."""

    def __init__(self, h0: float, omega: float, phase: float = 0.0, parent=None, time: float = 0.0):
        self.h0, self.omega, self.phase, self.time = float(h0), float(omega), float(phase), float(time)
        h = self.h0 * math.sin(self.omega * self.time + self.phase)
        super().__init__(offset=(0.0, 0.0, -h), parent=parent)

    def at_time(self, t: float) -> "HeaveFrame":
        return HeaveFrame(self.h0, self.omega, self.phase, self.parent.at_time(t) if self.parent else None, t)

    def save(self, group) -> None:
        for k in ("h0", "omega", "phase", "time"):
            group[k] = getattr(self, k)

    @classmethod
    def load(cls, group, parent=None) -> "HeaveFrame":
        return cls(float(group["h0"]), float(group["omega"]), float(group["phase"]), parent, float(group["time"]))

    def heave_rate(self) -> float:
        return self.h0 * self.omega * math.cos(self.omega * self.time + self.phase)

    def add_velocity(self, positions, velocity, /) -> None:
        super().add_velocity(positions, velocity)
        velocity[:, 2] += self.heave_rate()

    def device_terms(self):
        return KIND_TRANSLATING, np.zeros((3, 3)), np.array([0.0, 0.0, self.heave_rate()])


class PitchHeaveFrame(ReferenceFrame):
    """Pitch theta(t) = theta0 sin(omega t + phase_p) about the spanwise (y) axis through
    x = pivot_x, plus heave h(t) = h0 sin(omega t + phase_h) (no reference counterpart; the
    learned-start workload).  Body point p maps to P = Ry(theta) (p - pivot) + pivot + h z,
    where Ry(theta) (x, z) = (c x + s z, -s x + c z): the frame's angles are (0, -theta, 0)
    in the reference's convention (to_parent applies R^T) and its offset is chosen so that
    to_parent(p) = P.  ``add_velocity`` receives P and adds R (dP/dt), which the time
    step's following to_parent_without_offset turns into dP/dt."""

    def __init__(self, theta0: float, h0: float, omega: float, pivot_x: float = 0.25, phase_p: float = 0.0,
                 phase_h: float = 0.0, parent=None, time: float = 0.0):
        self.theta0, self.h0, self.omega = float(theta0), float(h0), float(omega)
        self.pivot_x, self.phase_p, self.phase_h, self.time = float(pivot_x), float(phase_p), float(phase_h), float(time)
        th, _, h, _ = self._state()
        c, s = math.cos(th), math.sin(th)
        # P = R^T (p - d) with R^T = Ry(theta): d = p - R P at p = 0, i.e. d = -R (pivot - Ry pivot + h z)
        qx = self.pivot_x - c * self.pivot_x
        qz = s * self.pivot_x + h
        # R = Ry(theta)^T: (x, z) -> (c x - s z, s x + c z)
        super().__init__(offset=(-(c * qx - s * qz), 0.0, -(s * qx + c * qz)), theta=(0.0, -th, 0.0), parent=parent)

    def _state(self):
        return self.motion(self.time)

    def motion(self, t: float) -> tuple[float, float, float, float]:
        """Pitch angle and rate, heave and heave rate at time t."""
        th = self.theta0 * math.sin(self.omega * t + self.phase_p)
        thd = self.theta0 * self.omega * math.cos(self.omega * t + self.phase_p)
        h = self.h0 * math.sin(self.omega * t + self.phase_h)
        hd = self.h0 * self.omega * math.cos(self.omega * t + self.phase_h)
        return th, thd, h, hd

    def save(self, group) -> None:
        for k in ("theta0", "h0", "omega", "pivot_x", "phase_p", "phase_h", "time"):
            group[k] = getattr(self, k)

    @classmethod
    def load(cls, group, parent=None) -> "PitchHeaveFrame":
        g = {k: float(group[k]) for k in ("theta0", "h0", "omega", "pivot_x", "phase_p", "phase_h", "time")}
        return cls(g["theta0"], g["h0"], g["omega"], g["pivot_x"], g["phase_p"], g["phase_h"], parent, g["time"])

    def at_time(self, t: float) -> "PitchHeaveFrame":
        return PitchHeaveFrame(self.theta0, self.h0, self.omega, self.pivot_x, self.phase_p, self.phase_h,
                               self.parent.at_time(t) if self.parent else None, t)

    def device_terms(self):
        # dP/dt = thd * (dRy/dtheta)(p - pivot) + hd z, with p - pivot = Ry^T (P - pivot - h z):
        # dP/dt = thd * J (P - pivot - h z) + hd z, J = [[0, 0, 1], [0, 0, 0], [-1, 0, 0]];
        # the frame adds R dP/dt, R = Ry^T.
        th, thd, h, hd = self._state()
        c, s = math.cos(th), math.sin(th)
        j = np.array([[0.0, 0.0, thd], [0.0, 0.0, 0.0], [-thd, 0.0, 0.0]])
        r = np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])
        w = r @ j
        base = np.array([self.pivot_x, 0.0, h])
        cvec = r @ (np.array([0.0, 0.0, hd]) - j @ base)
        return KIND_AFFINE, w, cvec

    def add_velocity(self, positions, velocity, /) -> None:
        super().add_velocity(positions, velocity)
        _, w, c = self.device_terms()
        velocity[:, :] += _vecmul_rows(w, positions) + c[None, :]


def device_chain_terms(rf: ReferenceFrame) -> tuple[np.ndarray, np.ndarray]:
    """The chain of an evaluated frame (child first) as the device kinematics takes it:
    24 doubles per frame (R^T, offset, W, c) and a kind per frame.  A frame class that
    overrides add_velocity must describe that velocity with device_terms."""
    terms, kinds = [], []
    f = rf
    while f is not None:
        cls = type(f)
        if cls.add_velocity is not ReferenceFrame.add_velocity and cls.device_terms is ReferenceFrame.device_terms:
            raise TypeError(f"{cls.__name__} overrides add_velocity without describing it to the device "
                            "(device_terms)")
        kind, w, c = f.device_terms()
        terms.append(np.concatenate([f.rotation_matrix_inverse.reshape(-1), f.offset, np.asarray(w).reshape(-1),
                                     np.asarray(c).reshape(-1)]))
        kinds.append(kind)
        f = f.parent
    if len(kinds) > 8:
        raise ValueError(f"frame chain of {len(kinds)} frames; at most 8 are supported")
    return np.ascontiguousarray(np.array(terms, np.float64).reshape(-1, 24)), np.array(kinds, np.int32)


def frame_chain_kinematics(rf: ReferenceFrame, positions, t: float) -> tuple[np.ndarray, np.ndarray]:
    """The time step's chain walk on the host (python/pyvl/solver.py:180-193): positions and
    velocities of a body's points in the global frame at time t."""
    new_rf = rf.at_time(t)
    pos = np.array(positions, np.float64)
    vel = np.zeros_like(pos)
    f = new_rf
    while f is not None:
        pos = f.to_parent_with_offset(pos, pos)
        f.add_velocity(pos, vel)
        vel = f.to_parent_without_offset(vel, vel)
        f = f.parent
    return pos, vel


def rf_to_serial(frame: ReferenceFrame):
    """A frame chain as a map: type, the frame's own save() data, and its parent
    (python/pyvl/geometry.py:65-76)."""
    from .fio import HirearchicalMap

    out = HirearchicalMap()
    out.insert_type("type", type(frame))
    data = HirearchicalMap()
    frame.save(data)
    out.insert_hirearchycal_map("data", data)
    if frame.parent is not None:
        out.insert_hirearchycal_map("parent", rf_to_serial(frame.parent))
    return out


def rf_from_serial(group) -> ReferenceFrame:
    """python/pyvl/geometry.py:79-88."""
    cls = group.get_type("type")
    parent = rf_from_serial(group.get_hirearchical_map("parent")) if "parent" in group else None
    return cls.load(group.get_hirearchical_map("data"), parent=parent)
