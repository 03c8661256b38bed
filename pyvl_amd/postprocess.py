"""Post-processing of solver results on the GPU's N-body kernel.

Mirrors python/pyvl/postprocess/{velocity,pressure,forces}.py: the same functions, arguments
and per-output-step loops over a :class:`pyvl_amd.solver.SolverResults`.  The reference
forms ``Mesh.induction_matrix`` (an M x N x 3 tensor) and contracts it with ``np.vecdot``;
here every induced velocity is the point-velocity sum (``pvl_point_velocity``: sum over lines
of line circulation times K, no tensor), the wake's velocity is its own sum on the device,
and the line forces run in ``pvl_line_forces``.  Free stream, density and the element-wise
combinations follow the reference's numpy expressions on the host.
"""

from __future__ import annotations

from collections.abc import Iterable

import numpy as np


def _points(p) -> np.ndarray:
    cpts = np.ascontiguousarray(p, dtype=np.float64)
    if len(cpts.shape) != 2 or cpts.shape[1] != 3:
        raise ValueError("Positions must be an array of 3 component position vectors.")
    return cpts


def _induced(results, i: int, t: float, cpts: np.ndarray) -> np.ndarray:
    pos = results.geometry.positions_at_time(t)
    return results.geometry.mesh.point_velocity(results.settings.model_settings.vortex_limit, pos,
                                                results.circulations[i, :], cpts)


def compute_velocities(results, positions) -> np.ndarray:
    """(n_out, N, 3) velocity at ``positions`` for every output step
    (python/pyvl/postprocess/velocity.py:11-46)."""
    out_times = results.settings.time_settings.output_times
    cpts = _points(positions)
    out = np.empty((out_times.size, cpts.shape[0], 3), np.float64)
    for i, t in enumerate(out_times):
        out[i] = _induced(results, i, t, cpts)
        out[i] += results.settings.flow_conditions.get_velocity(t, cpts)
        wm = results.wake_models[i]
        if wm is not None:
            out[i] += wm.get_velocity(cpts)
    return out


def compute_velocities_variable(results, positions: Iterable) -> list[np.ndarray]:
    """Velocity at a different point set per output step (velocity.py:49-89)."""
    out_times = results.settings.time_settings.output_times
    out: list[np.ndarray] = []
    for i, (t, pts) in enumerate(zip(out_times, positions, strict=True)):
        cpts = _points(pts)
        v = _induced(results, i, t, cpts)
        v += results.settings.flow_conditions.get_velocity(t, cpts)
        wm = results.wake_models[i]
        if wm is not None:
            v += wm.get_velocity(cpts)
        out.append(v)
    return out


def compute_surface_dynamic_pressure(results) -> list[np.ndarray]:
    """Dynamic pressure at the surface centres per output step (pressure.py:11-55):
    p = -rho u . (u/2 + V_inf), u the induced velocity relative to the moving surface."""
    out: list[np.ndarray] = []
    for i, t in enumerate(results.settings.time_settings.output_times):
        geo = results.geometry
        pos = geo.positions_at_time(t)
        cpts = geo.mesh.surface_average_vec3(pos)
        induced = geo.mesh.point_velocity(results.settings.model_settings.vortex_limit, pos,
                                          results.circulations[i, :], cpts)
        induced -= geo.mesh.surface_average_vec3(geo.velocity_at_time(t))
        fs = results.settings.flow_conditions.get_velocity(t, cpts)
        wm = results.wake_models[i]
        if wm is not None:
            induced += wm.get_velocity(cpts)
        p = np.vecdot(induced, 0.5 * induced + fs, axis=-1)
        out.append(-results.settings.flow_conditions.get_density(t, cpts) * p)
    return out


def compute_dynamic_pressure_variable(results, positions: Iterable) -> list[np.ndarray]:
    """Dynamic pressure at a different point set per output step (pressure.py:58-107)."""
    out: list[np.ndarray] = []
    for i, (t, pts) in enumerate(zip(results.settings.time_settings.output_times, positions, strict=True)):
        cpts = _points(pts)
        induced = _induced(results, i, t, cpts)
        fs = results.settings.flow_conditions.get_velocity(t, cpts)
        wm = results.wake_models[i]
        if wm is not None:
            induced += wm.get_velocity(cpts)
        p = np.vecdot(induced, 0.5 * induced + fs, axis=-1)
        out.append(-results.settings.flow_conditions.get_density(t, cpts) * p)
    return out


def circulatory_forces(results) -> list[np.ndarray]:
    """Forces on the mesh lines from the circulation, per output step
    (python/pyvl/postprocess/forces.py:10-42): reduced circulation Gamma / (2 pi) in the
    free stream + induced - motion velocity at the mesh nodes (Mesh.line_forces), then the
    wake's correction of the shedding lines."""
    from .cvl import Mesh

    out: list[np.ndarray] = []
    geo = results.geometry
    for i, t in enumerate(results.settings.time_settings.output_times):
        circ = results.circulations[i, :]
        reduced_c = circ / (2 * np.pi)
        positions = geo.positions_at_time(t)
        freestream = results.settings.flow_conditions.get_velocity(t, positions)
        induced = geo.mesh.point_velocity(results.settings.model_settings.vortex_limit, positions, circ, positions)
        motion = geo.velocity_at_time(t)
        wm = results.wake_models[i]
        if wm is not None:
            induced += wm.get_velocity(positions)
        forces = Mesh.line_forces(geo.mesh, geo.dual, np.ascontiguousarray(reduced_c), positions,
                                  np.ascontiguousarray(freestream + induced - motion))
        if wm is not None:
            wm.correct_forces(forces, geo, positions, circ, results.settings.flow_conditions)
        out.append(forces)
    return out
