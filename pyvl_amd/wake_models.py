"""``WakeModelLineExplicitUnsteady`` with its state in HBM.

Mirrors python/pyvl/wake_models.py:17-291 (constructor arguments, ``update``,
``get_velocity``, ``apply_corrections``, ``save``/``load`` and the ``wake_positions`` /
``circulation`` arrays in the reference layout).  Arguments may be numpy arrays (copied in
and out, as the reference's in-place numpy updates) or :class:`DeviceArray` (resident).

``mode="reference"`` reproduces the reference's convection (free stream + bound-mesh
induction, :121-128).  ``mode="free"`` also adds the wake's own induction (its
``get_velocity``, :159-175) to the convection velocity: a free wake.  The free stream is
evaluated per node on the GPU for the fields of :mod:`pyvl_amd.flow_conditions` (uniform and
rotating); any other object with ``get_velocity`` is taken as uniform, its velocity read at
the origin.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .cvl import INVALID_ID, Mesh
from .device import DeviceArray

_MODES = {"reference": 0, "free": 1}


def _flow_descriptor(flow, time: float) -> np.ndarray:
    if hasattr(flow, "device_descriptor"):
        return np.ascontiguousarray(flow.device_descriptor(), np.float64)
    v = np.asarray(flow.get_velocity(time, np.zeros((1, 3))), np.float64).reshape(-1, 3)[0]
    return np.array([0.0, v[0], v[1], v[2], 0.0, 0.0, 0.0], np.float64)


class WakeModelLineExplicitUnsteady:
    """Wake panels shed one row per step from the shedding lines."""

    def __init__(self, bordering_nodes, shedding_lines, adjacent_surfaces, vortex_tol: float, time: float,
                 line_rows: int, *, mode: str = "reference", precision: int = _lib.PVL_FP64) -> None:
        if len(shedding_lines) == 0:
            raise RuntimeError("The geometry has no edges which would shed any wake.")
        if mode not in _MODES:
            raise ValueError(f"mode must be one of {sorted(_MODES)}")
        self.shedding_lines = np.array(shedding_lines, np.uint)
        self.bordering_nodes = np.array(bordering_nodes, np.uint)
        self.adjacent_surfaces = np.array(adjacent_surfaces)
        self.line_rows = int(line_rows)
        self.vortex_tol = float(vortex_tol)
        self.mode = mode
        self.precision = int(precision)
        n_shed = self.shedding_lines.size
        bn = np.ascontiguousarray(self.bordering_nodes.reshape(n_shed, 2), np.uint32)
        adj = np.ascontiguousarray(self.adjacent_surfaces.reshape(n_shed, 2).astype(np.uint64) & 0xFFFFFFFF,
                                   np.uint32)
        h = C.c_void_p()
        _lib.call("pvl_wake_create", C.byref(h), C.c_uint(n_shed), C.c_uint(self.line_rows),
                  bn.ctypes.data_as(C.POINTER(C.c_uint32)), adj.ctypes.data_as(C.POINTER(C.c_uint32)),
                  C.c_double(self.vortex_tol), C.c_double(float(time)))
        self._h = h
        self._wake_mesh = None

    # -- state ----------------------------------------------------------------------------
    def _info(self):
        n_te, n_lines, steps = C.c_uint(), C.c_uint(), C.c_uint()
        t = C.c_double()
        _lib.call("pvl_wake_info", self._h, C.byref(n_te), C.byref(n_lines), C.byref(steps), C.byref(t))
        return n_te.value, n_lines.value, steps.value, t.value

    @property
    def step_count(self) -> int:
        return self._info()[2]

    @property
    def current_time(self) -> float:
        return self._info()[3]

    @property
    def n_unique_nodes(self) -> int:
        return self._info()[0] * (self.line_rows + 1)

    @property
    def n_unique_lines(self) -> int:
        return self._info()[1]

    def state(self) -> tuple[np.ndarray, np.ndarray]:
        n = self.shedding_lines.size
        pos = np.empty((n, 2, self.line_rows + 1, 3), np.float64)
        circ = np.empty((n, self.line_rows), np.float64)
        _lib.call("pvl_wake_get_state", self._h, pos.ctypes.data_as(C.c_void_p), circ.ctypes.data_as(C.c_void_p))
        return pos, circ

    @property
    def wake_positions(self) -> np.ndarray:
        """(n_shed, 2, line_rows + 1, 3) copy in the reference layout."""
        return self.state()[0]

    @property
    def circulation(self) -> np.ndarray:
        """(n_shed, line_rows) copy in the reference layout."""
        return self.state()[1]

    @property
    def wake_mesh(self) -> Mesh:
        """The reference's ring mesh (python/pyvl/wake_models.py:35-52)."""
        if self._wake_mesh is None:
            self._wake_mesh = Mesh.wake_rings(self.shedding_lines.size, self.line_rows)
        return self._wake_mesh

    # -- model ----------------------------------------------------------------------------
    def update(self, time: float, geometry, positions, circulation, flow) -> None:
        """Shed and convect one row (python/pyvl/wake_models.py:83-157)."""
        mesh = geometry.mesh if hasattr(geometry, "mesh") else geometry
        pos = positions if isinstance(positions, DeviceArray) else DeviceArray.from_numpy(
            np.ascontiguousarray(positions, np.float64))
        circ = circulation if isinstance(circulation, DeviceArray) else DeviceArray.from_numpy(
            np.ascontiguousarray(circulation, np.float64))
        desc = _flow_descriptor(flow, time)
        _lib.call("pvl_wake_update_flow", self._h, C.c_double(float(time)), mesh._device(), pos.cptr, circ.cptr,
                  desc.ctypes.data_as(C.POINTER(C.c_double)), C.c_int(_MODES[self.mode]), C.c_int(self.precision))

    def begin_update(self) -> None:
        """Start the next update's wake self-induction on a second stream (free wake only; no
        reference counterpart: the results of :meth:`update` are unchanged, it may only finish
        sooner because the work overlaps what is issued in between)."""
        _lib.call("pvl_wake_begin_update", self._h, C.c_int(_MODES[self.mode]), C.c_int(self.precision))

    def get_velocity(self, positions):
        """Velocity induced by the wake at ``positions`` (python/pyvl/wake_models.py:159-175)."""
        if isinstance(positions, DeviceArray):
            out = DeviceArray(positions.shape)
            _lib.call("pvl_wake_get_velocity", self._h, C.c_uint(positions.shape[0]), positions.cptr, out.cptr,
                      C.c_int(0), C.c_int(self.precision))
            return out
        p = DeviceArray.from_numpy(np.ascontiguousarray(positions, np.float64))
        return self.get_velocity(p).numpy()

    def apply_corrections(self, control_pts, normals, mat_in, rhs_in) -> None:
        """Implicit (matrix) and explicit (rhs) wake corrections, in place
        (python/pyvl/wake_models.py:177-223)."""
        dev = isinstance(mat_in, DeviceArray)
        cp = control_pts if isinstance(control_pts, DeviceArray) else DeviceArray.from_numpy(
            np.ascontiguousarray(control_pts, np.float64))
        nr = normals if isinstance(normals, DeviceArray) else DeviceArray.from_numpy(
            np.ascontiguousarray(normals, np.float64))
        a = mat_in if dev else DeviceArray.from_numpy(mat_in)
        b = rhs_in if isinstance(rhs_in, DeviceArray) else DeviceArray.from_numpy(rhs_in)
        _lib.call("pvl_wake_apply_corrections", self._h, C.c_uint(cp.shape[0]), cp.cptr, nr.cptr, a.cptr,
                  C.c_uint(a.shape[1]), b.cptr, C.c_int(self.precision))
        if not dev:
            a.numpy(mat_in)
        if not isinstance(rhs_in, DeviceArray):
            b.numpy(rhs_in)

    def correct_forces(self, line_forces, geometry, positions, circulation, flow) -> None:
        """Subtract the first wake row's force from the shedding lines' forces, in place
        (python/pyvl/wake_models.py:293-330): the row-0 rings as a one-row ring mesh with
        reduced circulation circulation[:, 0] / (2 pi), in the free stream plus the bound
        mesh's and the wake's induction at their nodes; shedding line i loses force 4 i + 1
        of that mesh (the reference's index).  The induction sums run on the GPU."""
        n = self.shedding_lines.size
        if getattr(self, "_row_mesh", None) is None:
            self._row_mesh = Mesh.wake_rings(n, 1)
            self._row_dual = self._row_mesh.compute_dual()
        wp, wc = self.state()
        cpts = np.ascontiguousarray(wp[:, :, 0:2, :].reshape(-1, 3))
        circ = np.ascontiguousarray(wc[:, 0] / (2 * np.pi))
        freestream = flow.get_velocity(self.current_time, cpts)
        mesh = geometry.mesh if hasattr(geometry, "mesh") else geometry
        ind_v = mesh.point_velocity(self.vortex_tol, np.ascontiguousarray(positions, np.float64),
                                    np.ascontiguousarray(circulation, np.float64), cpts)
        self_v = self.get_velocity(cpts)
        forces = Mesh.line_forces(self._row_mesh, self._row_dual, circ, cpts, freestream + ind_v + self_v)
        for i, i_line in enumerate(self.shedding_lines):
            line_forces[i_line, :] -= forces[4 * i + 1, :]

    # -- serialisation (python/pyvl/wake_models.py:233-291) ------------------------------
    def save(self) -> dict:
        pos, circ = self.state()
        return {
            "bordering_nodes": self.bordering_nodes.copy(),
            "adjacent_surfaces": self.adjacent_surfaces.copy(),
            "shedding_lines": self.shedding_lines.copy(),
            "current_time": self.current_time,
            "step_count": self.step_count,
            "line_rows": self.line_rows,
            "wake_positions": pos,
            "vortex_tol": self.vortex_tol,
            "circulation": circ,
        }

    @classmethod
    def load(cls, group: dict, **kw) -> "WakeModelLineExplicitUnsteady":
        out = cls(group["bordering_nodes"], group["shedding_lines"], group["adjacent_surfaces"],
                  float(group["vortex_tol"]), float(group["current_time"]), int(group["line_rows"]), **kw)
        out.set_state(group)
        return out

    def set_state(self, group) -> None:
        """Load positions, circulation, step count and time from a saved map (same topology)."""
        pos = np.ascontiguousarray(group["wake_positions"], np.float64)
        circ = np.ascontiguousarray(group["circulation"], np.float64)
        n = self.shedding_lines.size
        if pos.shape != (n, 2, self.line_rows + 1, 3) or circ.shape != (n, self.line_rows):
            raise ValueError("saved wake state does not match this wake's shape")
        _lib.call("pvl_wake_set_state", self._h, pos.ctypes.data_as(C.c_void_p), circ.ctypes.data_as(C.c_void_p),
                  C.c_uint(int(group["step_count"])), C.c_double(float(group["current_time"])))

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and _lib.alive():
                _lib._LIB.pvl_wake_destroy(self._h)
                self._h = None
        except Exception:
            pass


class WakeSnapshot:
    """A saved wake state (what the reference keeps per output step as
    ``type(wm).load(wm.save())``, python/pyvl/solver.py:253-254), held in host memory and
    evaluated on one shared device wake that is loaded with the state on use."""

    def __init__(self, saved: dict, evaluator: "WakeModelLineExplicitUnsteady"):
        self._saved = saved
        self._eval = evaluator

    def save(self) -> dict:
        return self._saved

    def __getattr__(self, name):
        if name in ("wake_positions", "circulation", "current_time", "step_count", "line_rows", "vortex_tol",
                    "shedding_lines", "bordering_nodes", "adjacent_surfaces"):
            return self._saved[name]
        raise AttributeError(name)

    def _loaded(self) -> "WakeModelLineExplicitUnsteady":
        ev = self._eval
        if getattr(ev, "_loaded_from", None) is not self:
            ev.set_state(self._saved)
            ev._loaded_from = self
        return ev

    def get_velocity(self, positions):
        return self._loaded().get_velocity(positions)

    def correct_forces(self, line_forces, geometry, positions, circulation, flow) -> None:
        self._loaded().correct_forces(line_forces, geometry, positions, circulation, flow)


def line_adjacency_information(mesh: Mesh, dual: Mesh, lines):
    """``SimulationGeometry.line_adjecency_information`` (python/pyvl/geometry.py:654-681):
    end nodes of each line and the surfaces on either side (dual line ends)."""
    lines = np.asarray(lines, np.int64)
    ld = mesh.line_data
    dd = dual.line_data
    bordering = np.stack([ld[lines, 0], ld[lines, 1]], axis=1).astype(np.uint) & INVALID_ID
    adjacent = np.stack([dd[lines, 0], dd[lines, 1]], axis=1).astype(np.uint) & INVALID_ID
    return bordering, adjacent
