"""The ``run_solver`` time step (python/pyvl/solver.py:145-265) kept resident on the GPU.

Per step, in the reference's order (solver.py:175-239):
  kinematics -> control points, control-point velocities, normals -> rhs
  -> system matrix (induction_matrix3) -> wake corrections (iteration > 0)
  -> LU solve -> closed-body mean removal -> wake update.

Kinematics plug in through the reference frames' ``at_time`` / ``add_velocity`` contract
(src/referenceframeobject.c:700-729, :939-955; :mod:`pyvl_amd.reference_frames`).  Each
step evaluates the frame chain at the step's time on the host (a few numbers per frame) and
walks it for every point on the GPU (``pvl_frame_kinematics``), so the body's positions and
velocities never leave HBM.
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
from .flow_conditions import FlowConditions, FlowConditionsRotating, FlowConditionsUniform  # noqa: F401
from .reference_frames import (HeaveFrame, PitchHeaveFrame, ReferenceFrame, RotorReferenceFrame,  # noqa: F401
                               TranslatingReferenceFrame, device_chain_terms)
from .settings import ModelSettings, SolverSettings, TimeSettings
from .wake_models import WakeModelLineExplicitUnsteady, WakeSnapshot, line_adjacency_information


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

    def __init__(self, mesh: Mesh, positions0: np.ndarray, frame, flow: FlowConditions, vortex_limit: float,
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
        self.d_p0 = DeviceArray.from_numpy(self.p0)
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
        self.flow_desc = np.ascontiguousarray(flow.device_descriptor(), np.float64)
        self.iteration = 0
        self.ir_iters: list[int] = []
        # "gmres": FP32 GMRES in FP64 refinement from warm_start(self) (numpy (n,) or None
        # for zero; by default the minimal-residual projection on the last three
        # circulations, learned.ProjectionStart) to max|b - A x| <= gmres_tol max|b|, LU64
        # fallback when it does not converge; gmres_stats: (corrections, GMRES steps, start
        # and final relative residual, converged) per solve
        self.warm_start = None
        if solve == "gmres":
            from .learned import ProjectionStart

            self.warm_start = ProjectionStart(n)
        self.gmres_tol = 1e-12
        self.gmres_stats: list[tuple] = []
        self.fallbacks = 0
        # start the free wake's self-induction at the beginning of each step, beside the
        # bound solve (bit-identical results; False: strictly in the reference's order)
        self.overlap_wake = True
        self.d_b0 = DeviceArray((n,)) if solve == "gmres" else None
        self.d_A32 = DeviceArray((n, n), np.float32) if solve == "gmres" else None

    def _kinematics(self, t: float) -> None:
        terms, kinds = device_chain_terms(self.frame.at_time(t))
        _lib.call("pvl_frame_kinematics", C.c_uint(self.p0.shape[0]), self.d_p0.cptr, C.c_uint(kinds.size),
                  terms.ctypes.data_as(C.POINTER(C.c_double)), kinds.ctypes.data_as(C.POINTER(C.c_int)),
                  self.d_pos.cptr, self.d_vel.cptr)
        _lib.call("pvl_surface_center", self.mesh._device(), self.d_pos.cptr, self.d_cp.cptr)
        _lib.call("pvl_surface_center", self.mesh._device(), self.d_vel.cptr, self.d_cpv.cptr)
        _lib.call("pvl_surface_normal", self.mesh._device(), self.d_pos.cptr, self.d_nrm.cptr)

    def assemble(self, t: float) -> None:
        """Kinematics, rhs, system matrix and wake corrections of one step."""
        self._kinematics(t)
        n = self.mesh.n_surfaces
        _lib.call("pvl_rhs_flow", C.c_uint(n), self.d_nrm.cptr, self.d_cpv.cptr, self.d_cp.cptr,
                  self.flow_desc.ctypes.data_as(C.POINTER(C.c_double)), self.d_b.cptr)
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

    # -- geometry, state snapshot and resume ------------------------------------------
    @property
    def geometry(self) -> "SimulationGeometry":
        if getattr(self, "_geometry", None) is None:
            self._geometry = SimulationGeometry(self.mesh, self.p0, self.frame, closed=self.closed)
        return self._geometry

    def snapshot(self) -> "SolverState":
        """The state after the last step (python/pyvl/solver.py:42-108: positions, normals,
        control points and their velocities, circulation, iteration, wake), copied to host
        memory; the device state is untouched."""
        st = SolverState(self.geometry, self.settings(), None)
        st.positions = self.d_pos.numpy()
        st.normals = self.d_nrm.numpy()
        st.control_points = self.d_cp.numpy()
        st.cp_velocity = self.d_cpv.numpy()
        st.circulation = self.d_x.numpy()
        st.iteration = self.iteration
        st.wake_data = self.wake.save() if self.wake is not None else None
        return st

    def restore(self, state: "SolverState") -> None:
        """Resume from a snapshot (no reference counterpart: the reference writes states but
        has no entry point that continues from one).  The next step is iteration
        ``state.iteration``: every later step is the one an uninterrupted run makes, bit for
        bit, since a step depends only on its time, the wake state and the geometry."""
        n = self.mesh.n_surfaces
        if state.circulation.shape != (n,):
            raise ValueError("state does not match this solver's geometry")
        self.iteration = int(state.iteration)
        self.d_pos.copy_from(np.ascontiguousarray(state.positions, np.float64))
        self.d_nrm.copy_from(np.ascontiguousarray(state.normals, np.float64))
        self.d_cp.copy_from(np.ascontiguousarray(state.control_points, np.float64))
        self.d_cpv.copy_from(np.ascontiguousarray(state.cp_velocity, np.float64))
        self.d_x.copy_from(np.ascontiguousarray(state.circulation, np.float64))
        if self.wake is not None:
            if state.wake_data is None:
                raise ValueError("state has no wake but this solver has one")
            self.wake.set_state(state.wake_data)

    def settings(self) -> SolverSettings:
        return SolverSettings(self.flow, ModelSettings(self.tol), TimeSettings(1, self.dt, None))


class SimulationGeometry:
    """One body's mesh, initial positions and reference frame: the single-body form of
    ``SimulationGeometry`` (python/pyvl/geometry.py:389-570) that the post-processing and
    the solver state read (``mesh``, ``dual``, counts, ``positions_at_time``,
    ``velocity_at_time``, save/load)."""

    def __init__(self, mesh: Mesh, positions, frame, closed: bool | None = None):
        self.mesh = mesh
        self.dual = mesh.compute_dual()
        self.positions = np.ascontiguousarray(positions, np.float64)
        self.rf = frame
        self.closed = is_closed(mesh) if closed is None else bool(closed)
        self.n_points, self.n_lines, self.n_surfaces = mesh.n_points, mesh.n_lines, mesh.n_surfaces

    def positions_at_time(self, t: float) -> np.ndarray:
        """python/pyvl/geometry.py:518-539: the evolved frame's to_global_with_offset."""
        return self.rf.at_time(float(t)).to_global_with_offset(self.positions)

    def velocity_at_time(self, t: float) -> np.ndarray:
        """python/pyvl/geometry.py:541-570: add_velocity at the current positions, then both
        taken to the parent, frame by frame."""
        rf = self.rf.at_time(float(t))
        pos = np.array(self.positions)
        v = np.zeros_like(pos)
        while rf is not None:
            rf.add_velocity(pos, v)
            rf.to_parent_with_offset(pos, pos)
            rf.to_parent_without_offset(v, v)
            rf = rf.parent
        return v

    def save(self):
        from .fio import HirearchicalMap
        from .reference_frames import rf_to_serial

        out = HirearchicalMap()
        mesh = HirearchicalMap()
        mesh.insert_int("n_points", self.mesh.n_points)
        counts, flat = self.mesh.to_element_connectivity()
        mesh.insert_array("n_per_element", counts)
        mesh.insert_array("flattened_elements", flat)
        out.insert_hirearchycal_map("mesh", mesh)
        out.insert_array("positions", self.positions)
        out.insert_hirearchycal_map("reference_frame", rf_to_serial(self.rf))
        out.insert_int("closed", int(self.closed))
        return out

    @classmethod
    def load(cls, group) -> "SimulationGeometry":
        from .reference_frames import rf_from_serial

        m = group.get_hirearchical_map("mesh")
        counts = m.get_array("n_per_element")
        flat = np.asarray(m.get_array("flattened_elements"), np.uint32)
        offs = np.pad(np.cumsum(counts), (1, 0))
        mesh = Mesh(m.get_int("n_points"), [flat[offs[i]:offs[i + 1]] for i in range(counts.size)])
        return cls(mesh, group.get_array("positions"), rf_from_serial(group.get_hirearchical_map("reference_frame")),
                   closed=bool(group.get_int("closed")))


class SolverState:
    """python/pyvl/solver.py:42-108: the step's arrays, the iteration, the settings, the
    geometry and the wake's saved state."""

    def __init__(self, geometry: SimulationGeometry, settings: SolverSettings, wake_data: dict | None):
        n, p = geometry.n_surfaces, geometry.n_points
        self.geometry = geometry
        self.settings = settings
        self.positions = np.empty((p, 3))
        self.normals = np.empty((n, 3))
        self.control_points = np.empty((n, 3))
        self.cp_velocity = np.empty((n, 3))
        self.circulation = np.empty((n,))
        self.iteration = 0
        self.wake_data = wake_data

    def save(self):
        from .fio import HirearchicalMap

        out = HirearchicalMap()
        for k in ("positions", "normals", "control_points", "cp_velocity", "circulation"):
            out.insert_array(k, getattr(self, k))
        out.insert_int("iteration", int(self.iteration))
        out.insert_hirearchycal_map("solver_settings", HirearchicalMap.from_dict(self.settings.save()))
        out.insert_hirearchycal_map("simulation_geometry", self.geometry.save())
        if self.wake_data is not None:
            wm = HirearchicalMap()
            wm.insert_string("type", "WakeModelLineExplicitUnsteady")
            wm.insert_hirearchycal_map("data", HirearchicalMap.from_dict(self.wake_data))
            out.insert_hirearchycal_map("wake_model", wm)
        return out

    @classmethod
    def load(cls, hmap) -> "SolverState":
        geometry = SimulationGeometry.load(hmap.get_hirearchical_map("simulation_geometry"))
        settings = SolverSettings.load(hmap.get_hirearchical_map("solver_settings"))
        wake = None
        if "wake_model" in hmap:
            wake = dict(hmap.get_hirearchical_map("wake_model").get_hirearchical_map("data").to_dict())
        st = cls(geometry, settings, wake)
        st.iteration = hmap.get_int("iteration")
        for k in ("positions", "normals", "control_points", "cp_velocity", "circulation"):
            setattr(st, k, hmap.get_array(k))
        return st


class SolverResults:
    """python/pyvl/solver.py:20-39: circulations at the output times and the wake's state at
    each (host snapshots evaluated on one shared device wake)."""

    def __init__(self, geometry: SimulationGeometry, settings: SolverSettings):
        self.geometry = geometry
        self.settings = settings
        self.wake_models: list = []
        self.circulations = np.empty((settings.time_settings.output_times.size, geometry.n_surfaces), np.float64)


@dataclass(frozen=True)
class OutputSettings:
    """File output of the state at each output step (python/pyvl/solver.py:114-142): "JSON"
    (the reference's JSON layout) or "NPZ" (exact binary; in place of the reference's HDF5)."""

    ftype: str
    naming_callback: object

    def serialization_fn(self, hmap, path) -> None:
        from . import fio

        if self.ftype == "JSON":
            fio.serialize_json(hmap, path)
        elif self.ftype == "NPZ":
            fio.serialize_npz(hmap, path)
        else:
            raise ValueError(f"The file type ftype={self.ftype!r} is not valid.")


def run_solver(solver: UnsteadySolver, time_settings: TimeSettings,
               output_settings: OutputSettings | None = None) -> SolverResults:
    """run_solver (python/pyvl/solver.py:145-265) over ``time_settings.nt`` steps of the
    resident solver, from its current iteration: circulations and wake snapshots at the
    output steps (every ``output_interval``), the state written through ``output_settings``."""
    if float(time_settings.dt) != solver.dt:
        raise ValueError(f"time step {time_settings.dt} differs from the solver's {solver.dt}")
    settings = SolverSettings(solver.flow, ModelSettings(solver.tol), time_settings)
    results = SolverResults(solver.geometry, settings)
    evaluator = None
    i_out = 0
    for iteration in range(time_settings.nt):
        solver.step()
        it = solver.iteration - 1
        if time_settings.output_interval is None or time_settings.output_interval == 0 or \
                iteration % time_settings.output_interval == 0:
            results.circulations[i_out, :] = solver.d_x.numpy()
            if solver.wake is not None:
                if evaluator is None:
                    w = solver.wake
                    evaluator = WakeModelLineExplicitUnsteady(w.bordering_nodes, w.shedding_lines,
                                                              w.adjacent_surfaces, w.vortex_tol, 0.0, w.line_rows,
                                                              mode=w.mode, precision=w.precision)
                results.wake_models.append(WakeSnapshot(solver.wake.save(), evaluator))
            else:
                results.wake_models.append(None)
            if output_settings is not None:
                output_settings.serialization_fn(solver.snapshot().save(),
                                                 output_settings.naming_callback(it, it * solver.dt))
            i_out += 1
    return results


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
