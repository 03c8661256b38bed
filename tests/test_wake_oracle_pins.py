"""Pins of oracle/wake_oracle.py (the restatement of python/pyvl/wake_models.py:83-223 and
python/pyvl/solver.py:175-239) by what the mathematics fixes, on the CPU.

The induction routines underneath are pinned bit-for-bit to the reference's own C core
(tests/test_oracle.py); these tests pin the wake bookkeeping built on them:

1. convection closed form: with zero bound circulation every wake node moves with the free
   stream, so after k updates row r holds the trailing-edge position of update
   max(k - r, 1) advanced r times (rows beyond the filled ones collapsed onto the last
   filled row) -- fixes the reversed row loop, the row shift and the collapse;
2. convection velocity: each advected row moves by (V_inf + bound-induced velocity) dt,
   the induced velocity taken by a second route (normal-projected assembly along the three
   axes instead of the vector tensor) -- fixes which positions and which sign;
3. implicit/explicit consistency: the corrected system A' G = b' is the uncorrected system
   with the full wake's normal velocity added, where the wake's newest row carries the
   circulation that the next update assigns to it (G[adj_1] - G[adj_0], pinned by the
   reference's own tests, test/pytests/test_wake_models.py:75-77,155-157) -- fixes the
   signs of the implicit columns;
4. correction split: the rhs correction equals the wake's normal velocity with row 0's
   circulation zeroed, and the matrix correction equals -/+ the normal induction of a
   separately built one-row ring wake -- fixes the reshape and the row split;
5. mirror symmetry: a y-symmetric wing with y-symmetric motion sheds a y-symmetric wake.

Each check is also run on a copy of the oracle with a planted mistake, which it must
reject.
"""

import copy

import numpy as np
import pytest

from oracle import oracle as O
from oracle import wake_oracle as WO
from pyvl_amd import geometry as G

INVALID = WO.INVALID


class _Heave:
    def __init__(self, h0, omega):
        self.h0, self.omega = h0, omega

    def positions(self, p0, t):
        out = np.array(p0, np.float64)
        out[:, 2] += self.h0 * np.sin(self.omega * t)
        return out

    def velocities(self, p0, t):
        v = np.zeros_like(p0)
        v[:, 2] = self.h0 * self.omega * np.cos(self.omega * t)
        return v


def _wing(n_span=8, n_chord=3, taper=0.6):
    pos, el, _ = G.flat_wing(n_span, n_chord, 4.0, 1.0, taper=taper)
    t = O.mesh_from_elements(pos.shape[0], el)
    dual_lines = O.mesh_dual(t)[0]
    ny = n_span + 1
    te_nodes = set(range(n_chord * ny, (n_chord + 1) * ny))
    lines = t.lines & O.MASK
    shed = np.array([i for i in range(t.n_lines) if int(lines[i, 0]) in te_nodes and int(lines[i, 1]) in te_nodes])
    nodes = lines[shed].astype(np.uint)
    adj = (dual_lines[shed] & O.MASK).astype(np.uint)
    return pos, t, shed, nodes, adj


def _run(wake_cls, steps, R=5, vinf=(1.0, 0.0, 0.08), dt=0.125, frame=None, wing=None):
    pos, t, shed, nodes, adj = wing or _wing()
    frame = frame or _Heave(0.1, 0.9)
    w = wake_cls(nodes, shed, adj, 1e-6, 0.0, R)
    vinf = np.asarray(vinf, np.float64)
    gammas = []
    for it in range(steps):
        g, _, _ = WO.step(it, dt, t, pos, frame, vinf, 1e-6, w)
        gammas.append(g)
    return pos, t, w, frame, vinf, gammas


# --------------------------------------------------------------------------------------
# planted mistakes (each one a plausible misreading of python/pyvl/wake_models.py)
class _ForwardOrder(WO.OracleWake):
    """convects rows first to last instead of last to first (:121)."""

    def update(self, time, bound, positions, circulation, vinf):
        dt = time - self.current_time
        self.current_time = time
        if self.step_count == 0:
            self.wake_positions[:, 0, 0, :] = positions[self.bordering_nodes[:, 0], :]
            self.wake_positions[:, 1, 0, :] = positions[self.bordering_nodes[:, 1], :]
            self.step_count = 1
        complete_rows = min(self.step_count, self.line_rows)
        for i in range(complete_rows):
            old_pos = self.wake_positions[:, :, i, :]
            flat = np.array(old_pos.reshape((-1, 3)), np.float64)
            vel = np.full_like(flat, vinf)
            ind = O.induction_matrix(bound, positions, flat, self.vortex_tol)
            vel += np.vecdot(ind, circulation[None, :, None], axis=1)
            self.wake_positions[:, :, i + 1, :] = old_pos + vel.reshape(old_pos.shape) * dt
        self.wake_positions[:, 0, 0, :] = positions[self.bordering_nodes[:, 0], :]
        self.wake_positions[:, 1, 0, :] = positions[self.bordering_nodes[:, 1], :]
        if complete_rows != self.line_rows:
            self.wake_positions[:, :, complete_rows + 1:, :] = self.wake_positions[:, :, complete_rows, :][:, :, None]
        self.circulation[:, 1:] = self.circulation[:, :-1]
        self.circulation[:, 0] = 0
        mask = self.adjacent_surfaces != INVALID
        self.circulation[mask[:, 1], 0] += circulation[self.adjacent_surfaces[mask[:, 1], 1]]
        self.circulation[mask[:, 0], 0] -= circulation[self.adjacent_surfaces[mask[:, 0], 0]]
        self.step_count += 1


class _InducedSignFlipped(WO.OracleWake):
    """subtracts the bound-induced velocity (:125-126)."""

    def update(self, time, bound, positions, circulation, vinf):
        super().update(time, bound, positions, -np.asarray(circulation), vinf)
        mask = self.adjacent_surfaces != INVALID
        self.circulation[:, 0] = 0
        self.circulation[mask[:, 1], 0] += circulation[self.adjacent_surfaces[mask[:, 1], 1]]
        self.circulation[mask[:, 0], 0] -= circulation[self.adjacent_surfaces[mask[:, 0], 0]]


class _ImplicitSignsSwapped(WO.OracleWake):
    """adds the implicit correction to column adj_1 and subtracts it from adj_0 (:207-214)."""

    def apply_corrections(self, control_pts, normals, mat_in, rhs_in):
        saved = self.adjacent_surfaces
        self.adjacent_surfaces = saved[:, ::-1].copy()
        try:
            super().apply_corrections(control_pts, normals, mat_in, rhs_in)
        finally:
            self.adjacent_surfaces = saved


class _ExplicitRowsShifted(WO.OracleWake):
    """pairs ring rows 1.. with circulation rows 0.. (:218-222)."""

    def apply_corrections(self, control_pts, normals, mat_in, rhs_in):
        saved = self.circulation
        shifted = np.zeros_like(saved)
        shifted[:, 1:] = saved[:, :-1]
        self.circulation = shifted
        try:
            super().apply_corrections(control_pts, normals, mat_in, rhs_in)
        finally:
            self.circulation = saved


# --------------------------------------------------------------------------------------
# 1. convection closed form
def _closed_form_error(wake_cls):
    pos, t, shed, nodes, adj = _wing()
    R = 5
    w = wake_cls(nodes, shed, adj, 1e-6, -0.1, R)
    frame = _Heave(0.1, 0.9)
    vinf = np.array([1.0, -0.03, 0.08])
    zero = np.zeros(t.n_surfaces)
    times = [0.0, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875, 1.0]
    te = {}
    dts = {}
    prev = -0.1
    worst = 0
    for k, tk in enumerate(times, start=1):
        p = frame.positions(pos, tk)
        te[k] = np.stack([p[nodes[:, 0]], p[nodes[:, 1]]], axis=1)  # (n_shed, 2, 3)
        dts[k] = tk - prev
        prev = tk
        w.update(tk, t, p, zero, vinf)
        want = np.empty_like(w.wake_positions)
        want[:, :, 0] = te[k]
        for r in range(1, R + 1):
            if r > k:
                want[:, :, r] = want[:, :, k]
                continue
            x = te[max(k - r, 1)].copy()
            for u in range(k - r + 1, k + 1):
                x = x + np.full_like(x, vinf) * dts[u]
            want[:, :, r] = x
        worst = max(worst, int(np.count_nonzero(want != w.wake_positions)))
    return worst


def test_convection_closed_form_bit_exact():
    assert _closed_form_error(WO.OracleWake) == 0


def test_convection_closed_form_rejects_forward_order():
    assert _closed_form_error(_ForwardOrder) > 0


# --------------------------------------------------------------------------------------
# 2. convection velocity by a second route
def _velocity_route_error(wake_cls):
    pos, t, w, frame, vinf, gammas = _run(wake_cls, 4)
    dt = 0.125
    it = 4
    tk = it * dt
    p = frame.positions(pos, tk)
    gamma = np.random.default_rng(5).normal(size=t.n_surfaces)
    before = w.wake_positions.copy()
    w.update(tk, t, p, gamma, vinf)
    after = w.wake_positions
    rows = min(it + 1, w.line_rows)  # step_count before this update was it + 1
    worst = 0.0
    axes = np.eye(3)
    for i in range(rows):
        pts = before[:, :, i, :].reshape(-1, 3)
        u = np.stack([O.induction_matrix3(t, p, pts, np.tile(axes[a], (pts.shape[0], 1)), 1e-6) @ gamma
                      for a in range(3)], axis=1)
        vel = (after[:, :, i + 1, :].reshape(-1, 3) - pts) / dt
        want = vinf[None, :] + u
        worst = max(worst, float(np.max(np.abs(vel - want)) / np.max(np.abs(want))))
    return worst


def test_convection_velocity_is_free_stream_plus_bound_induction():
    assert _velocity_route_error(WO.OracleWake) < 1e-12


def test_convection_velocity_rejects_flipped_induction():
    assert _velocity_route_error(_InducedSignFlipped) > 1e-6


# --------------------------------------------------------------------------------------
# 3./4. corrections
def _corrected_system(wake_cls, steps=6):
    pos, t, w, frame, vinf, gammas = _run(wake_cls, steps)
    dt = 0.125
    tk = steps * dt
    p = frame.positions(pos, tk)
    cps = O.surface_center(t, p)
    cpv = O.surface_center(t, frame.velocities(pos, tk))
    nrm = O.surface_normal(t, p)
    b = np.vecdot(nrm, -(np.full_like(cps, vinf) - cpv), axis=1)
    a = O.induction_matrix3(t, p, cps, nrm, 1e-6)
    ac, bc = a.copy(), b.copy()
    w.apply_corrections(cps, nrm, ac, bc)
    return t, w, cps, nrm, a, b, ac, bc


def _wake_normal_velocity(w, cps, nrm, circulation):
    ind = O.induction_matrix(w.wake_mesh, w.wake_positions.reshape(-1, 3), cps, w.vortex_tol)
    v = np.vecdot(ind, circulation.reshape(1, -1, 1), axis=1)
    return np.vecdot(nrm, v, axis=1)


def _consistency_error(wake_cls):
    t, w, cps, nrm, a, b, ac, bc = _corrected_system(wake_cls)
    gamma = np.linalg.solve(ac, bc)
    adj = w.adjacent_surfaces
    g_new = np.zeros(adj.shape[0])
    m1 = adj[:, 1] != INVALID
    m0 = adj[:, 0] != INVALID
    g_new[m1] += gamma[adj[m1, 1]]
    g_new[m0] -= gamma[adj[m0, 0]]
    circ = w.circulation.copy()
    circ[:, 0] = g_new
    lhs = a @ gamma - b
    rhs = _wake_normal_velocity(w, cps, nrm, circ)
    return float(np.max(np.abs(lhs - rhs)) / np.max(np.abs(b)))


def test_corrections_consistent_with_next_wake_row():
    assert _consistency_error(WO.OracleWake) < 1e-12


def test_corrections_consistency_rejects_swapped_implicit_signs():
    assert _consistency_error(_ImplicitSignsSwapped) > 1e-6


def _split_errors(wake_cls):
    t, w, cps, nrm, a, b, ac, bc = _corrected_system(wake_cls)
    circ = w.circulation.copy()
    circ[:, 0] = 0.0
    want_b = b + _wake_normal_velocity(w, cps, nrm, circ)
    err_b = float(np.max(np.abs(bc - want_b)) / np.max(np.abs(bc - b)))
    # one-row ring wake built on its own from the wake's first two node rows
    one = WO.OracleWake(w.bordering_nodes, w.shedding_lines, w.adjacent_surfaces, w.vortex_tol, 0.0, 1)
    one.wake_positions[...] = w.wake_positions[:, :, 0:2, :]
    col = O.induction_matrix3(one.wake_mesh, one.wake_positions.reshape(-1, 3), cps, nrm, w.vortex_tol)
    want_a = a.copy()
    for i, (s0, s1) in enumerate(w.adjacent_surfaces):
        if s1 != INVALID:
            want_a[:, s1] -= col[:, i]
        if s0 != INVALID:
            want_a[:, s0] += col[:, i]
    err_a = float(np.max(np.abs(ac - want_a)) / np.max(np.abs(ac - a)))
    return err_a, err_b


def test_correction_split_rows_and_columns():
    err_a, err_b = _split_errors(WO.OracleWake)
    assert err_a < 1e-12 and err_b < 1e-12


def test_correction_split_rejects_shifted_explicit_rows():
    err_a, err_b = _split_errors(_ExplicitRowsShifted)
    assert err_b > 1e-6


# --------------------------------------------------------------------------------------
# 5. mirror symmetry
def _mirror_error(wake_positions):
    pts = wake_positions.reshape(-1, 3)
    mir = pts * np.array([1.0, -1.0, 1.0])
    d = np.linalg.norm(pts[:, None, :] - mir[None, :, :], axis=2)
    return float(np.max(np.min(d, axis=1)) / np.max(np.abs(pts)))


def test_mirror_symmetric_wake():
    pos, t, w, frame, vinf, gammas = _run(WO.OracleWake, 7, vinf=(1.0, 0.0, 0.08))
    assert _mirror_error(w.wake_positions) < 1e-13
    # strips are in span order and every ring is numbered by the same pattern, so strip i
    # and its mirror image n-1-i carry the same ring circulation (as mirrored panels do)
    c = w.circulation
    assert np.max(np.abs(c)) > 0
    assert np.max(np.abs(c - c[::-1])) <= 1e-12 * np.max(np.abs(c))


def test_mirror_check_rejects_asymmetric_motion():
    """The symmetry check is not vacuous: a side-slip breaks it."""
    pos, t, w, frame, vinf, gammas = _run(WO.OracleWake, 7, vinf=(1.0, 0.05, 0.08))
    assert _mirror_error(w.wake_positions) > 1e-6


def test_step_zero_has_no_wake_corrections():
    """python/pyvl/solver.py:217: corrections start at iteration 1, the wake's first update
    pins row 0 to the trailing edge (:114-117)."""
    pos, t, shed, nodes, adj = _wing()
    frame = _Heave(0.1, 0.9)
    w = WO.OracleWake(nodes, shed, adj, 1e-6, 0.0, 4)
    w.wake_positions[...] = np.nan  # the reference's np.empty: never read at iteration 0
    g, a, b = WO.step(0, 0.125, t, pos, frame, np.array([1.0, 0.0, 0.08]), 1e-6, w)
    assert np.all(np.isfinite(g))
    p = frame.positions(pos, 0.0)
    assert np.array_equal(w.wake_positions[:, 0, 0], p[nodes[:, 0]])
    assert np.array_equal(w.wake_positions[:, 1, 0], p[nodes[:, 1]])
    assert np.allclose(g, np.linalg.solve(a, b), rtol=1e-12, atol=0)
    w2 = copy.deepcopy(w)
    g1, _, _ = WO.step(1, 0.125, t, pos, frame, np.array([1.0, 0.0, 0.08]), 1e-6, w2)
    assert not np.allclose(g1, g, rtol=1e-3)
