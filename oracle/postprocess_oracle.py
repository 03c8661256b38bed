"""ORACLE -- test infrastructure only.  Restatement of the reference's post-processing for
one output step, on the oracle's C induction (oracle/pvl_oracle.c) and wake
(oracle/wake_oracle.py):

  velocity_at      python/pyvl/postprocess/velocity.py:34-46
  surface_pressure python/pyvl/postprocess/pressure.py:29-53
  pressure_at      python/pyvl/postprocess/pressure.py:80-106
  forces           python/pyvl/postprocess/forces.py:13-41 with Mesh.line_forces
                   (src/meshobject.c:1350-1393) and
                   WakeModelLineExplicitUnsteady.correct_forces (python/pyvl/wake_models.py:293-330)

Arguments are one output step's data: the bound topology, positions and point velocities
(positions_at_time / velocity_at_time), circulation, uniform free stream, density, vortex
limit and the wake's oracle state.  The induction sums are the reference's
np.vecdot(induction_matrix, circulation) taken by O.point_velocity (the same sum without the
tensor).  Mesh.line_forces is pinned bit for bit against the reference's own C
(tests/test_postprocess_oracle.py); the compositions follow the cited lines.
"""

from __future__ import annotations

import numpy as np

from . import oracle as O
from . import wake_oracle as WO


def velocity_at(topo, pos, circ, pts, vinf, tol, wake=None):
    out = O.point_velocity(topo, pos, circ, pts, tol)
    out += np.full_like(pts, vinf)
    if wake is not None:
        out += wake.get_velocity(pts)
    return out


def pressure_at(topo, pos, circ, pts, vinf, rho, tol, wake=None):
    induced = O.point_velocity(topo, pos, circ, pts, tol)
    fs = np.full_like(pts, vinf)
    if wake is not None:
        induced += wake.get_velocity(pts)
    return -rho * np.vecdot(induced, 0.5 * induced + fs, axis=-1)


def surface_pressure(topo, pos, vel, circ, vinf, rho, tol, wake=None):
    cpts = O.surface_center(topo, pos)
    induced = O.point_velocity(topo, pos, circ, cpts, tol)
    induced -= O.surface_center(topo, vel)
    fs = np.full_like(cpts, vinf)
    if wake is not None:
        induced += wake.get_velocity(cpts)
    return -rho * np.vecdot(induced, 0.5 * induced + fs, axis=-1)


def correct_forces(wake: WO.OracleWake, line_forces, topo, pos, circ, vinf):
    n = wake.shedding_lines.size
    tmp = WO.wake_ring_topology(n, 1)
    dual_lines, _, _ = O.mesh_dual(tmp)
    cpts = np.ascontiguousarray(wake.wake_positions[:, :, 0:2, :].reshape(-1, 3))
    c = np.ascontiguousarray(wake.circulation[:, 0] / (2 * np.pi))
    fs = np.full_like(cpts, vinf)
    ind_v = O.point_velocity(topo, pos, circ, cpts, wake.vortex_tol)
    self_v = wake.get_velocity(cpts)
    f = O.line_forces(tmp, dual_lines, c, cpts, fs + ind_v + self_v)
    for i, i_line in enumerate(wake.shedding_lines):
        line_forces[i_line, :] -= f[4 * i + 1, :]


def forces(topo, pos, vel, circ, vinf, tol, wake=None):
    dual_lines, _, _ = O.mesh_dual(topo)
    reduced = circ / (2 * np.pi)
    fs = np.full_like(pos, vinf)
    induced = O.point_velocity(topo, pos, circ, pos, tol)
    if wake is not None:
        induced += wake.get_velocity(pos)
    f = O.line_forces(topo, dual_lines, reduced, pos, fs + induced - vel)
    if wake is not None:
        correct_forces(wake, f, topo, pos, circ, vinf)
    return f
