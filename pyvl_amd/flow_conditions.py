"""Free-stream fields: ``FlowConditionsUniform`` and ``FlowConditionsRotating``
(python/pyvl/flow_conditions.py:143-541), with the reference's fields, defaults,
``get_velocity`` / ``get_density`` / ``get_pressure`` and save/load maps.

Each field also describes itself to the device (:meth:`FlowConditions.device_descriptor`,
seven doubles {kind, a[3], b[3]}): the time step's right-hand side (``pvl_rhs_flow``) and the
wake's convection (``pvl_wake_update_flow``) evaluate it per point on the GPU.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


class FlowConditions:
    """Base class: velocity, density and static pressure fields."""

    def get_velocity(self, time, positions, out_array=None):  # pragma: no cover - interface
        raise NotImplementedError

    def get_density(self, time, positions, out_array=None):
        del time
        if out_array is not None:
            out_array[...] = self.rho
            return out_array
        return np.full(np.asarray(positions).shape[0], self.rho)

    def get_pressure(self, time, positions, out_array=None):
        del time
        if out_array is not None:
            out_array[...] = self.p_stat
            return out_array
        return np.full(np.asarray(positions).shape[0], self.p_stat)

    def device_descriptor(self) -> np.ndarray:  # pragma: no cover - interface
        raise NotImplementedError


@dataclass(frozen=True)
class FlowConditionsUniform(FlowConditions):
    """Uniform velocity (vx, vy, vz) (python/pyvl/flow_conditions.py:143-330)."""

    vx: float
    vy: float
    vz: float
    rho: float = 1.225
    p_stat: float = 101325

    def get_velocity(self, time, positions, out_array=None):
        del time
        v = np.array((self.vx, self.vy, self.vz), np.float64)
        if out_array is not None:
            out_array[:, :] = v[None, :]
            return out_array
        return np.full_like(np.asarray(positions, np.float64), v)

    @property
    def vector(self) -> np.ndarray:
        return np.array((self.vx, self.vy, self.vz), np.float64)

    def device_descriptor(self) -> np.ndarray:
        return np.array([0.0, self.vx, self.vy, self.vz, 0.0, 0.0, 0.0], np.float64)

    def save(self) -> dict:
        return {"type": "FlowConditionsUniform", "velocity": np.array((self.vx, self.vy, self.vz), np.float64)}

    @classmethod
    def load(cls, group: dict) -> "FlowConditionsUniform":
        return cls(*np.asarray(group["velocity"], np.float64))


@dataclass(frozen=True)
class FlowConditionsRotating(FlowConditions):
    """Rotation about a centre r0 with angular velocity w: v(r) = (r - r0) x w
    (python/pyvl/flow_conditions.py:332-541)."""

    center_x: float
    center_y: float
    center_z: float
    omega_x: float
    omega_y: float
    omega_z: float
    rho: float = 1.225
    p_stat: float = 101325.0

    def get_velocity(self, time, positions, out_array=None):
        """:412-446; like the reference, ``out_array`` is not written."""
        del time, out_array
        pos = np.asarray(positions, np.float64) - np.array(((self.center_x, self.center_y, self.center_z),),
                                                           np.float64)
        omg = np.array((self.omega_x, self.omega_y, self.omega_z), np.float64)
        return np.cross(pos, omg)

    def device_descriptor(self) -> np.ndarray:
        return np.array([1.0, self.center_x, self.center_y, self.center_z, self.omega_x, self.omega_y,
                         self.omega_z], np.float64)

    def save(self) -> dict:
        return {"type": "FlowConditionsRotating",
                "center": np.array((self.center_x, self.center_y, self.center_z), np.float64),
                "omega": np.array((self.omega_x, self.omega_y, self.omega_z), np.float64)}

    @classmethod
    def load(cls, group: dict) -> "FlowConditionsRotating":
        return cls(*np.asarray(group["center"], np.float64), *np.asarray(group["omega"], np.float64))


_TYPES = {"FlowConditionsUniform": FlowConditionsUniform, "FlowConditionsRotating": FlowConditionsRotating}


def load_flow_conditions(group: dict) -> FlowConditions:
    """Rebuild a saved field from its map (the type name is stored with it)."""
    return _TYPES[str(group["type"])].load(group)
