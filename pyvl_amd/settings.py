"""Solver settings: ``TimeSettings``, ``ModelSettings``, ``SolverSettings``
(python/pyvl/settings.py), with the same fields, ``output_times`` rule and save/load maps."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class TimeSettings:
    """``nt`` steps of ``dt``; output every ``output_interval`` steps (every step if None or 0)."""

    nt: int
    dt: float
    output_interval: int | None = None

    @property
    def simulation_times(self) -> np.ndarray:
        return np.arange(self.nt, dtype=np.float64) * np.float64(self.dt)

    @property
    def output_times(self) -> np.ndarray:
        """python/pyvl/settings.py:39-45."""
        if self.output_interval is None or self.output_interval == 0:
            return self.simulation_times
        return np.float64(self.dt) * np.arange(self.nt, step=self.output_interval, dtype=np.float64)

    def save(self) -> dict:
        out = {"nt": int(self.nt), "dt": float(self.dt)}
        if self.output_interval is not None:
            out["output_interval"] = int(self.output_interval)
        return out

    @classmethod
    def load(cls, group: dict) -> "TimeSettings":
        return cls(nt=int(group["nt"]), dt=float(group["dt"]), output_interval=group.get("output_interval"))


@dataclass
class ModelSettings:
    """``vortex_limit``: the cut-off distance below which a line induces nothing."""

    vortex_limit: float

    def save(self) -> dict:
        return {"vortex_limit": float(self.vortex_limit)}

    @classmethod
    def load(cls, group: dict) -> "ModelSettings":
        return cls(vortex_limit=float(group["vortex_limit"]))


@dataclass(frozen=True)
class SolverSettings:
    """Flow conditions, model settings and time settings of a run."""

    flow_conditions: object
    model_settings: ModelSettings
    time_settings: TimeSettings = TimeSettings(1, 1, None)

    def save(self) -> dict:
        return {"flow_conditions": self.flow_conditions.save(), "model_settings": self.model_settings.save(),
                "time_settings": self.time_settings.save()}

    @classmethod
    def load(cls, group: dict, flow_types: dict | None = None) -> "SolverSettings":
        from .flow_conditions import load_flow_conditions

        return cls(flow_conditions=load_flow_conditions(group["flow_conditions"]),
                   model_settings=ModelSettings.load(group["model_settings"]),
                   time_settings=TimeSettings.load(group["time_settings"]))
