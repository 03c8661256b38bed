"""The ``run_solver`` time step (python/pyvl/solver.py:145-265) kept resident on the GPU.

Per step, in the reference's order (solver.py:175-239):
  kinematics -> control points, control-point velocities, normals -> rhs
  -> system matrix (induction_matrix3) -> wake corrections (iteration > 0)
  -> LU solve -> closed-body mean removal -> wake update.

Kinematics plug in through the reference frames' ``at_time`` / ``add_velocity`` contract
(src/referenceframeobject.c:700-729, :939-955); :class:`HeaveFrame` and
:class:`PitchHeaveFrame` are closed-form frames for the oscillating-wing workloads (the
reference has translating and rotor frames only).
"""

from __future__ import annotations

import ctypes as C
import time as _time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import linalg as LA
from .cvl import INVALID_ID, Mesh
from .device import DeviceArray
from .wake_models import WakeModelLineExplicitUnsteady, line_adjacency_information


@dataclass(frozen=True)
class FlowConditionsUniform:
    """Uniform free stream (python/pyvl/flow_conditions.py:201-233)."""

    vx: float
    vy: float
    vz: float

    def get_velocity(self, time, positions, out_array=None):
        v = np.array((self.vx, self.vy, self.vz), np.float64)
        if out_array is not None:
            out_array[:, :] = v[None, :]
            return out_array
        return np.full_like(np.asarray(positions, np.float64), v)

    @property
    def vector(self) -> np.ndarray:
        return np.array((self.vx, self.vy, self.vz), np.float64)


@dataclass(frozen=True)
class StaticFrame:
    """Identity frame (``ReferenceFrame()``)."""

    def positions(self, p0: np.ndarray, t: float) -> np.ndarray:
        return p0

    def velocities(self, p0: np.ndarray, t: float) -> np.ndarray:
        return np.zeros_like(p0)


@dataclass(frozen=True)
class HeaveFrame:
    """Vertical oscillation z(t) = h0 sin(omega t + phase)."""

    h0: float
    omega: float
    phase: float = 0.0

    def positions(self, p0, t):
        out = np.array(p0, np.float64)
        out[:, 2] += self.h0 * np.sin(self.omega * t + self.phase)
        return out

    def velocities(self, p0, t):
        v = np.zeros_like(p0)
        v[:, 2] = self.h0 * self.omega * np.cos(self.omega * t + self.phase)
        return v


@dataclass(frozen=True)
class PitchHeaveFrame:
    """Pitch theta(t) = theta0 sin(omega t + phase_p) about the spanwise (y) axis through
    ``pivot_x``, plus heave h(t) = h0 sin(omega t + phase_h)."""

    theta0: float
    h0: float
    omega: float
    pivot_x: float = 0.25
    phase_p: float = 0.0
    phase_h: float = 0.0

    def _angles(self, t):
        th = self.theta0 * np.sin(self.omega * t + self.phase_p)
        thd = self.theta0 * self.omega * np.cos(self.omega * t + self.phase_p)
        h = self.h0 * np.sin(self.omega * t + self.phase_h)
        hd = self.h0 * self.omega * np.cos(self.omega * t + self.phase_h)
        return th, thd, h, hd

    def positions(self, p0, t):
        th, _, h, _ = self._angles(t)
        c, s = np.cos(th), np.sin(th)
        x = p0[:, 0] - self.pivot_x
        z = p0[:, 2]
        out = np.array(p0, np.float64)
        out[:, 0] = self.pivot_x + c * x + s * z
        out[:, 2] = -s * x + c * z + h
        return out

    def velocities(self, p0, t):
        th, thd, _, hd = self._angles(t)
        c, s = np.cos(th), np.sin(th)
        x = p0[:, 0] - self.pivot_x
        z = p0[:, 2]
        v = np.zeros_like(p0)
        v[:, 0] = thd * (-s * x + c * z)
        v[:, 2] = thd * (-c * x - s * z) + hd
        return v


def is_closed(mesh: Mesh) -> bool:
    """python/pyvl/geometry.py:463-469: closed when no dual line has an INVALID_ID end."""
    d = mesh.compute_dual().line_data
    return not bool(np.any((d & INVALID_ID) == INVALID_ID))


def trailing_edge_lines(mesh: Mesh, positions: np.ndarray, x_te: float, tol: float = 1e-12) -> np.ndarray:
    """Lines with both ends at x = x_te (examples/plot_example_2.py:63-67)."""
    ld = mesh.line_data
    x = positions[:, 0]
    sel = (np.abs(x[ld[:, 0]] - x_te) <= tol) & (np.abs(x[ld[:, 1]] - x_te) <= tol)
    return np.nonzero(sel)[0]


@dataclass
class StepTiming:
    wall_s: list = field(default_factory=list)


class UnsteadySolver:
    """One body (optionally with a wake), every array of the step resident in HBM.

    ``solve``: "lu64" (FP64 LU, the reference's arithmetic) or "ir32" (FP32 LU on MFMA +
    FP64 iterative refinement to the relative residual ``ir_tol``).
    """

    def __init__(self, mesh: Mesh, positions0: np.ndarray, frame, flow: FlowConditionsUniform, vortex_limit: float,
                 dt: float, wake: WakeModelLineExplicitUnsteady | None = None, *, induction_precision: int = 0,
                 solve: str = "lu64", ir_tol: float = 1e-13, closed: bool | None = None):
        self.mesh = mesh
        self.p0 = np.ascontiguousarray(positions0, np.float64)
        self.frame = frame
        self.flow = flow
        self.tol = float(vortex_limit)
        self.dt = float(dt)
        self.wake = wake
        self.prec = int(induction_precision)
        self.solve = solve
        self.ir_tol = ir_tol
        self.closed = is_closed(mesh) if closed is None else closed
        n, p = mesh.n_surfaces, mesh.n_points
        self.d_pos = DeviceArray((p, 3))
        self.d_vel = DeviceArray((p, 3))
        self.d_cp = DeviceArray((n, 3))
        self.d_cpv = DeviceArray((n, 3))
        self.d_nrm = DeviceArray((n, 3))
        self.d_A = DeviceArray((n, n))
        self.d_b = DeviceArray((n,))
        self.d_x = DeviceArray((n,))
        # partial pivoting's factorisation; without interchanges when it makes none
        self.lu = LA.LU(n, LA.FP32 if solve == "ir32" else LA.FP64, pivoting="auto")
        self.vinf = np.ascontiguousarray(flow.vector)
        self.iteration = 0
        self.ir_iters: list[int] = []
        # "gmres": FP32 GMRES in FP64 refinement from warm_start(self) (numpy (n,) or None
        # for zero) to max|b - A x| <= gmres_tol max|b|, LU64 fallback when it does not
        # converge; gmres_stats: (corrections, GMRES steps, start and final relative
        # residual, converged) per solve
        self.warm_start = None
        self.gmres_tol = 1e-12
        self.gmres_stats: list[tuple] = []
        self.fallbacks = 0
        # start the free wake's self-induction at the beginning of each step, beside the
        # bound solve (bit-identical results; False: strictly in the reference's order)
        self.overlap_wake = True
        self.d_b0 = DeviceArray((n,)) if solve == "gmres" else None
        self.d_A32 = DeviceArray((n, n), np.float32) if solve == "gmres" else None

    def _kinematics(self, t: float) -> None:
        self.d_pos.copy_from(self.frame.positions(self.p0, t))
        self.d_vel.copy_from(self.frame.velocities(self.p0, t))
        _lib.call("pvl_surface_center", self.mesh._device(), self.d_pos.cptr, self.d_cp.cptr)
        _lib.call("pvl_surface_center", self.mesh._device(), self.d_vel.cptr, self.d_cpv.cptr)
        _lib.call("pvl_surface_normal", self.mesh._device(), self.d_pos.cptr, self.d_nrm.cptr)

    def assemble(self, t: float) -> None:
        """Kinematics, rhs, system matrix and wake corrections of one step."""
        self._kinematics(t)
        n = self.mesh.n_surfaces
        _lib.call("pvl_rhs_normal_velocity", C.c_uint(n), self.d_nrm.cptr, self.d_cpv.cptr,
                  self.vinf.ctypes.data_as(C.POINTER(C.c_double)), self.d_b.cptr)
        self.mesh.induction_matrix3(self.tol, self.d_pos, self.d_cp, self.d_nrm, self.d_A, precision=self.prec)
        if self.d_b0 is not None:
            _lib.call("pvl_memcpy_d2d", self.d_b0.cptr, self.d_b.cptr, C.c_size_t(8 * n))
        if self.wake is not None and self.iteration > 0:
            self.wake.apply_corrections(self.d_cp, self.d_nrm, self.d_A, self.d_b)

    def _lu64(self) -> None:
        n = self.mesh.n_surfaces
        self.lu.factor(self.d_A)
        _lib.call("pvl_memcpy_d2d", self.d_x.cptr, self.d_b.cptr, C.c_size_t(8 * n))
        self.lu.solve(self.d_x)

    def solve_system(self) -> None:
        n = self.mesh.n_surfaces
        if self.solve == "gmres":
            from .learned import gmres_refine

            _lib.call("pvl_convert_f64_f32", self.d_A.cptr, self.d_A32.cptr, C.c_size_t(n * n))
            x0 = self.warm_start(self) if self.warm_start is not None else None
            if x0 is None:
                _lib.call("pvl_memset_zero", self.d_x.cptr, C.c_size_t(8 * n))
            elif isinstance(x0, DeviceArray):
                _lib.call("pvl_memcpy_d2d", self.d_x.cptr, x0.cptr, C.c_size_t(8 * n))
            else:
                self.d_x.copy_from(np.ascontiguousarray(x0, np.float64))
            ok, outer, steps, rel0, rel1 = gmres_refine(self.d_A, self.d_A32, self.d_b, self.d_x, tol=self.gmres_tol)
            self.gmres_stats.append((outer, steps, rel0, rel1, ok))
            if not ok:
                self.fallbacks += 1
                self._lu64()
        elif self.solve == "lu64":
            self.lu.factor(self.d_A)
            _lib.call("pvl_memcpy_d2d", self.d_x.cptr, self.d_b.cptr, C.c_size_t(8 * n))
            self.lu.solve(self.d_x)
        else:
            self.lu.factor(self.d_A)
            _, it, _, _ = LA.solve_refine(self.d_A, self.d_b, self.lu, tol=self.ir_tol, out=self.d_x)
            self.ir_iters.append(it)
        if self.closed:
            _lib.call("pvl_remove_mean", self.d_x.cptr, C.c_int(0), C.c_int(n))

    def step(self) -> DeviceArray:
        t = self.iteration * self.dt
        if self.overlap_wake and self.wake is not None:
            # the wake's self-induction needs only the wake state: it runs beside the solve
            self.wake.begin_update()
        self.assemble(t)
        self.solve_system()
        if self.wake is not None:
            self.wake.update(t, self.mesh, self.d_pos, self.d_x, self.flow)
        self.iteration += 1
        return self.d_x

    def run(self, n_steps: int, callback=None) -> list[float]:
        """Run n_steps; returns per-step wall times (s), each ending in a device sync."""
        walls = []
        for _ in range(n_steps):
            t0 = _time.perf_counter()
            self.step()
            _lib.call("pvl_synchronize")
            walls.append(_time.perf_counter() - t0)
            if callback is not None:
                callback(self)
        return walls


def make_wing_wake_solver(n_span: int, n_chord: int, span: float, chord: float, frame, flow, line_rows: int,
                          dt: float, *, vortex_limit: float = 1e-6, taper: float = 1.0, mode: str = "reference",
                          induction_precision: int = 0, solve: str = "lu64") -> "UnsteadySolver":
    """Flat (optionally tapered) wing shedding from its trailing edge, as in
    examples/plot_example_2.py:51-90 with the wake model of python/pyvl/wake_models.py."""
    from .geometry import flat_wing

    pos0, elements, _ = flat_wing(n_span, n_chord, span, chord, taper=taper)
    mesh = Mesh(pos0.shape[0], elements)
    dual = mesh.compute_dual()
    te_x = pos0[:, 0]
    # trailing edge: the chord-end node row
    ny = n_span + 1
    te_nodes = set(range(n_chord * ny, (n_chord + 1) * ny))
    ld = mesh.line_data
    shed = np.array([i for i in range(mesh.n_lines) if int(ld[i, 0]) in te_nodes and int(ld[i, 1]) in te_nodes])
    del te_x
    nodes, surfs = line_adjacency_information(mesh, dual, shed)
    wake = WakeModelLineExplicitUnsteady(nodes, shed, surfs, vortex_limit, 0.0, line_rows, mode=mode,
                                         precision=induction_precision)
    solver = UnsteadySolver(mesh, pos0, frame, flow, vortex_limit, dt, wake, induction_precision=induction_precision,
                            solve=solve, closed=False)
    return solver


def c5_kinematics(seed: int) -> "PitchHeaveFrame":
    """Seeded random pitch + heave: pitch amplitude U(0, 10 deg), heave amplitude U(0, 0.2)
    chord, reduced frequency k U(0.05, 0.5) (omega = 2 k U / c with U = c = 1), phases
    U(0, 2 pi), pitch axis at quarter chord."""
    rng = np.random.default_rng(seed)
    theta0 = np.radians(rng.uniform(0.0, 10.0))
    h0 = rng.uniform(0.0, 0.2)
    k = rng.uniform(0.05, 0.5)
    return PitchHeaveFrame(theta0, h0, 2.0 * k, 0.25, rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi))


C5 = dict(n_span=64, n_chord=32, span=8.0, chord=1.0, taper=0.5, rows=300, dt=1.0 / 32.0, tol=1e-6)


def make_c5_solver(kin_seed: int, solve: str = "lu64", mode: str = "free", rows: int | None = None):
    """Pitching + heaving tapered wing, 64 x 32 = 2,048 panels (span 8, root chord 1, taper
    0.5), free wake of ``rows`` rows (300 by default), FP64 induction."""
    fc = FlowConditionsUniform(1.0, 0.0, 0.0)
    return make_wing_wake_solver(C5["n_span"], C5["n_chord"], C5["span"], C5["chord"], c5_kinematics(kin_seed), fc,
                                 rows or C5["rows"], C5["dt"], vortex_limit=C5["tol"], taper=C5["taper"], mode=mode,
                                 solve=solve)
