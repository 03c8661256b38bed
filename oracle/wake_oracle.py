"""ORACLE -- test infrastructure only.  numpy restatement of the reference's unsteady wake
and time step, on the oracle's C induction (oracle/pvl_oracle.c) and scipy's LAPACK LU (the
solver the reference calls).

Follows python/pyvl/wake_models.py:35-223 and python/pyvl/solver.py:175-239 line by line,
in the reference layout: wake_positions (n_shed, 2, R+1, 3), circulation (n_shed, R), the
full ring mesh of _create_wake_mesh with every (also unfilled) row evaluated each step, and
the per-row convection loop.  Pinned by the reference tests' circulation shift
(test/pytests/test_wake_models.py:75-77,155-157) and by what the mathematics fixes
(tests/test_wake_oracle_pins.py): the free-stream convection closed form (bit-exact), the
convection velocity by a second induction route, the consistency of the implicit columns
with the next update's newest-row circulation, the row split of the corrections, and
mirror symmetry; each check also rejects a copy with a planted mistake.
"""

from __future__ import annotations

import numpy as np
import scipy.linalg as la

from . import frames_oracle as FO
from . import oracle as O

INVALID = 0x7FFFFFFF


def free_stream(vinf, t, pts):
    """The free stream at points: ``vinf`` is a uniform (3,) velocity or a field
    ``vinf(t, pts)`` (flow.get_velocity, python/pyvl/wake_models.py:124)."""
    if callable(vinf):
        return np.asarray(vinf(t, pts), np.float64)
    return np.full_like(pts, vinf)


def rotating_flow(center, omega):
    """FlowConditionsRotating.get_velocity (python/pyvl/flow_conditions.py:412-446):
    v = (p - center) x omega."""
    c = np.asarray(center, np.float64)[None, :]
    w = np.asarray(omega, np.float64)

    def field(t, pts):
        return np.cross(np.asarray(pts, np.float64) - c, w)

    return field


def wake_ring_topology(n_lines: int, line_rows: int) -> O.Topo:
    """_create_wake_mesh (python/pyvl/wake_models.py:35-52) through mesh_from_elements."""
    npr = line_rows + 1
    ring = np.empty((n_lines, line_rows, 4), np.uint32)
    ring[:, :, 0] = np.arange(line_rows)
    ring[:, :, 1] = np.arange(line_rows) + npr
    ring[:, :, 2] = np.arange(line_rows) + npr + 1
    ring[:, :, 3] = np.arange(line_rows) + 1
    ring += (2 * npr * np.arange(n_lines, dtype=np.uint32))[:, None, None]
    return O.mesh_from_elements(2 * npr * n_lines, ring.reshape(-1, 4).tolist())


class OracleWake:
    """WakeModelLineExplicitUnsteady restated (python/pyvl/wake_models.py:54-223)."""

    def __init__(self, bordering_nodes, shedding_lines, adjacent_surfaces, vortex_tol, time, line_rows):
        self.shedding_lines = np.array(shedding_lines, np.uint)
        n = self.shedding_lines.size
        self.wake_positions = np.zeros((n, 2, line_rows + 1, 3), np.float64)
        self.line_rows = line_rows
        self.bordering_nodes = np.array(bordering_nodes, np.uint)
        self.adjacent_surfaces = np.array(adjacent_surfaces)
        self.vortex_tol = vortex_tol
        self.circulation = np.zeros((n, line_rows), np.float64)
        self.wake_mesh = wake_ring_topology(n, line_rows)
        self.current_time = time
        self.step_count = 0

    def update(self, time, bound: O.Topo, positions, circulation, vinf):
        dt = time - self.current_time
        self.current_time = time
        if self.step_count == 0:
            self.wake_positions[:, 0, 0, :] = positions[self.bordering_nodes[:, 0], :]
            self.wake_positions[:, 1, 0, :] = positions[self.bordering_nodes[:, 1], :]
            self.step_count = 1
        complete_rows = min(self.step_count, self.line_rows)
        for i in reversed(range(complete_rows)):
            old_pos = self.wake_positions[:, :, i, :]
            flat = np.array(old_pos.reshape((-1, 3)), np.float64)
            vel = free_stream(vinf, time, flat)
            ind_mat = O.induction_matrix(bound, positions, flat, self.vortex_tol)
            vel += np.vecdot(ind_mat, circulation[None, :, None], axis=1)
            new_pos = old_pos + vel.reshape(old_pos.shape) * dt
            self.wake_positions[:, :, i + 1, :] = new_pos
        self.wake_positions[:, 0, 0, :] = positions[self.bordering_nodes[:, 0], :]
        self.wake_positions[:, 1, 0, :] = positions[self.bordering_nodes[:, 1], :]
        if complete_rows != self.line_rows:
            self.wake_positions[:, :, complete_rows + 1:, :] = (self.wake_positions[:, :, complete_rows, :])[:, :, None, :]
        self.circulation[:, 1:] = self.circulation[:, :-1]
        self.circulation[:, 0] = 0
        mask = self.adjacent_surfaces != INVALID
        self.circulation[mask[:, 1], 0] += circulation[self.adjacent_surfaces[mask[:, 1], 1]]
        self.circulation[mask[:, 0], 0] -= circulation[self.adjacent_surfaces[mask[:, 0], 0]]
        self.step_count += 1

    def get_velocity(self, positions):
        ind = O.induction_matrix(self.wake_mesh, self.wake_positions.reshape((-1, 3)), positions, self.vortex_tol)
        return np.vecdot(ind, self.circulation.reshape((1, -1, 1)), axis=1)

    def apply_corrections(self, control_pts, normals, mat_in, rhs_in):
        n = self.shedding_lines.size
        ind_mat = O.induction_matrix3(self.wake_mesh, self.wake_positions.reshape((-1, 3)), control_pts, normals,
                                      self.vortex_tol).reshape((control_pts.shape[0], n, self.line_rows))
        implicit = ind_mat[:, :, 0]
        explicit = ind_mat[:, :, 1:].reshape((control_pts.shape[0], -1))
        for i_line in range(n):
            correction = implicit[:, i_line]
            s1 = self.adjacent_surfaces[i_line, 1]
            if s1 != INVALID:
                mat_in[:, s1] -= correction
            s2 = self.adjacent_surfaces[i_line, 0]
            if s2 != INVALID:
                mat_in[:, s2] += correction
        circ = self.circulation.reshape((n, self.line_rows))[:, 1:].reshape((-1,))
        rhs_in[:] += np.vecdot(explicit, circ[None, :], axis=1)


def free_update(ow: OracleWake, time, bound: O.Topo, positions, circulation, vinf):
    """Free-wake update as a composition of reference functions: every filled node's velocity
    is the free stream + the bound mesh's induction (python/pyvl/wake_models.py:124-126) +
    the wake's own get_velocity (:159-175), all from the state at the start of the update;
    then the reference's update bookkeeping (:114-155).  Point-velocity sums go through
    O.point_velocity (the tensor contraction without the tensor)."""
    dt = time - ow.current_time
    if ow.step_count == 0:
        ow.wake_positions[:, 0, 0, :] = positions[ow.bordering_nodes[:, 0], :]
        ow.wake_positions[:, 1, 0, :] = positions[ow.bordering_nodes[:, 1], :]
        ow.step_count = 1
    cr = min(ow.step_count, ow.line_rows)
    old = ow.wake_positions[:, :, :cr, :].copy()
    flat = np.ascontiguousarray(old.reshape(-1, 3))
    vel = free_stream(vinf, time, flat)
    vel += O.point_velocity(bound, positions, circulation, flat, ow.vortex_tol)
    vel += O.point_velocity(ow.wake_mesh, ow.wake_positions.reshape(-1, 3), ow.circulation.reshape(-1), flat,
                            ow.vortex_tol)
    ow.wake_positions[:, :, 1:cr + 1, :] = old + vel.reshape(old.shape) * dt
    ow.current_time = time
    ow.wake_positions[:, 0, 0, :] = positions[ow.bordering_nodes[:, 0], :]
    ow.wake_positions[:, 1, 0, :] = positions[ow.bordering_nodes[:, 1], :]
    if cr != ow.line_rows:
        ow.wake_positions[:, :, cr + 1:, :] = ow.wake_positions[:, :, cr, :][:, :, None, :]
    ow.circulation[:, 1:] = ow.circulation[:, :-1]
    ow.circulation[:, 0] = 0
    mask = ow.adjacent_surfaces != INVALID
    ow.circulation[mask[:, 1], 0] += circulation[ow.adjacent_surfaces[mask[:, 1], 1]]
    ow.circulation[mask[:, 0], 0] -= circulation[ow.adjacent_surfaces[mask[:, 0], 0]]
    ow.step_count += 1


def reference_core_step(it: int, dt: float, bound: O.Topo, p0, frame, vinf, tol, ow: OracleWake):
    """One iteration of python/pyvl/solver.py:175-239 with the wake, every induction taken
    from the reference's own C core (oracle/_ref: compute_line_induction and the two line-
    buffer reductions, as src/meshobject.c calls them) and the reference's numpy steps
    restated: the reference path itself, for checking at full size.  Returns the
    circulation."""
    t = it * dt
    pos, vel = FO.kinematics(frame, p0, t)
    cps = O.ref_surface_center(bound, pos)
    cpv = O.ref_surface_center(bound, vel)
    nrm = O.ref_surface_normal(bound, pos)
    b = np.vecdot(nrm, -(free_stream(vinf, t, cps) - cpv), axis=1)
    a = O.ref_induction_matrix3(bound, pos, cps, nrm, tol)
    if it > 0:
        n = ow.shedding_lines.size
        ind = O.ref_induction_matrix3(ow.wake_mesh, ow.wake_positions.reshape(-1, 3), cps, nrm, tol)
        ind = ind.reshape(cps.shape[0], n, ow.line_rows)
        for i_line in range(n):
            corr = ind[:, i_line, 0]
            s1, s0 = ow.adjacent_surfaces[i_line, 1], ow.adjacent_surfaces[i_line, 0]
            if s1 != INVALID:
                a[:, s1] -= corr
            if s0 != INVALID:
                a[:, s0] += corr
        b += np.vecdot(ind[:, :, 1:].reshape(cps.shape[0], -1), ow.circulation[:, 1:].reshape(-1)[None, :], axis=1)
    gamma = la.lu_solve(la.lu_factor(a), b)
    dt_w = t - ow.current_time
    ow.current_time = t
    if ow.step_count == 0:
        ow.wake_positions[:, 0, 0, :] = pos[ow.bordering_nodes[:, 0], :]
        ow.wake_positions[:, 1, 0, :] = pos[ow.bordering_nodes[:, 1], :]
        ow.step_count = 1
    cr = min(ow.step_count, ow.line_rows)
    for i in reversed(range(cr)):
        old = ow.wake_positions[:, :, i, :]
        fl = np.array(old.reshape(-1, 3))
        v = free_stream(vinf, t, fl)
        v += np.vecdot(O.ref_induction_matrix(bound, pos, fl, tol), gamma[None, :, None], axis=1)
        ow.wake_positions[:, :, i + 1, :] = old + v.reshape(old.shape) * dt_w
    ow.wake_positions[:, 0, 0, :] = pos[ow.bordering_nodes[:, 0], :]
    ow.wake_positions[:, 1, 0, :] = pos[ow.bordering_nodes[:, 1], :]
    if cr != ow.line_rows:
        ow.wake_positions[:, :, cr + 1:, :] = ow.wake_positions[:, :, cr, :][:, :, None, :]
    ow.circulation[:, 1:] = ow.circulation[:, :-1]
    ow.circulation[:, 0] = 0
    m = ow.adjacent_surfaces != INVALID
    ow.circulation[m[:, 1], 0] += gamma[ow.adjacent_surfaces[m[:, 1], 1]]
    ow.circulation[m[:, 0], 0] -= gamma[ow.adjacent_surfaces[m[:, 0], 0]]
    ow.step_count += 1
    return gamma


def step(it: int, dt: float, bound: O.Topo, p0, frame, vinf, tol, wake: OracleWake | None, closed=False,
         free=False):
    """One iteration of python/pyvl/solver.py:175-239 for a single body (free=True: the
    wake convects with free_update)."""
    t = it * dt
    pos, vel = FO.kinematics(frame, p0, t)
    cps = O.surface_center(bound, pos)
    cpv = O.surface_center(bound, vel)
    nrm = O.surface_normal(bound, pos)
    ev = free_stream(vinf, t, cps) - cpv
    b = np.vecdot(nrm, -ev, axis=1)
    a = O.induction_matrix3(bound, pos, cps, nrm, tol)
    if wake is not None and it > 0:
        wake.apply_corrections(cps, nrm, a, b)
    gamma = la.lu_solve(la.lu_factor(a), b)
    if closed:
        gamma -= np.mean(gamma)
    if wake is not None:
        if free:
            free_update(wake, t, bound, pos, gamma, vinf)
        else:
            wake.update(t, bound, pos, gamma, vinf)
    return gamma, a, b
