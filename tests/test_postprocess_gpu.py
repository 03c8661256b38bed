"""GPU paths added around the time step, each against the oracle:

- Mesh.line_forces (pvl_line_forces): bit-equal to the oracle's restatement, which equals
  the reference's own C (tests/test_postprocess_oracle.py);
- rigid-frame kinematics on the device (pvl_frame_kinematics): bit-equal to the oracle's
  chain walk for the reference's frame kinds, within 4 ulp-scale of the closed forms of the
  oscillating frames;
- post-processing (pyvl_amd.postprocess) of a solver run: velocities, dynamic pressures and
  circulatory forces with the wake's correction, per output step, against
  oracle/postprocess_oracle.py on the same circulations and wake states (<= 1e-12);
- non-uniform free stream (FlowConditionsRotating) in the time loop, and a rotor frame:
  against the oracle's loop (no teacher forcing, <= 1e-10);
- state snapshot and resume through the JSON and binary files: bit-identical continuation.
"""

import ctypes as C

import numpy as np
import pytest

from oracle import frames_oracle as FO
from oracle import oracle as O
from oracle import postprocess_oracle as PO
from oracle import wake_oracle as WO
from pyvl_amd import _lib, fio
from pyvl_amd import geometry as G
from pyvl_amd import postprocess as PP
from pyvl_amd import reference_frames as RF
from pyvl_amd.cvl import Mesh
from pyvl_amd.device import DeviceArray
from pyvl_amd.solver import (FlowConditionsRotating, FlowConditionsUniform, SolverState, TimeSettings,
                             make_wing_wake_solver, run_solver)

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(31)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_line_forces_kernel_is_bit_exact():
    for pos, el in (G.flat_wing(16, 6, 4.0, 1.0, taper=0.5)[:2], G.uv_sphere(32, 16, sigma=1e-3, seed=2)):
        pos = pos + 0.01 * RNG.normal(size=pos.shape)
        m = Mesh(pos.shape[0], el)
        d = m.compute_dual()
        circ = RNG.normal(size=m.n_surfaces)
        vel = RNG.normal(size=pos.shape)
        got = Mesh.line_forces(m, d, circ, pos, vel)
        topo = O.mesh_from_elements(pos.shape[0], el)
        want = O.line_forces(topo, O.mesh_dual(topo)[0], circ, pos, vel)
        assert np.array_equal(got, want)
        out = np.empty_like(got)
        assert Mesh.line_forces(m, d, circ, pos, vel, out) is out and np.array_equal(out, want)
    with pytest.raises(ValueError):
        Mesh.line_forces(m, m, circ, pos, vel)


def _device_kinematics(frame, p0, t):
    terms, kinds = RF.device_chain_terms(frame.at_time(t))
    dp0 = DeviceArray.from_numpy(p0)
    pos, vel = DeviceArray(p0.shape), DeviceArray(p0.shape)
    _lib.call("pvl_frame_kinematics", C.c_uint(p0.shape[0]), dp0.cptr, C.c_uint(kinds.size),
              terms.ctypes.data_as(C.POINTER(C.c_double)), kinds.ctypes.data_as(C.POINTER(C.c_int)), pos.cptr,
              vel.cptr)
    return pos.numpy(), vel.numpy()


def _mixed_chain():
    root = RF.RotorReferenceFrame((0.5, 0.0, -1.0), (0.1, 0.0, 0.2), None, (0.0, 0.2, 1.5), 0.0)
    mid = RF.TranslatingReferenceFrame((0.0, 1.0, 0.0), (0.0, 0.5, 0.0), (1.0, 0.0, 0.25), root, 0.25)
    return RF.ReferenceFrame((0.3, -0.2, 0.1), (1.0, 2.0, 3.0), mid)


@pytest.mark.parametrize("t", [0.0, 0.45, 3.0])
def test_frame_kinematics_kernel_matches_oracle(t):
    p0 = RNG.normal(size=(5000, 3))
    pos, vel = _device_kinematics(_mixed_chain(), p0, t)
    po, vo = FO.kinematics(_mixed_chain(), p0, t)
    assert np.array_equal(pos, po) and np.array_equal(vel, vo)
    for f in (RF.HeaveFrame(0.1, 0.6, 0.2), RF.PitchHeaveFrame(np.radians(8.0), 0.12, 0.8, 0.25, 0.3, 1.1)):
        pos, vel = _device_kinematics(f, p0, t)
        po, vo = FO.kinematics(f, p0, t)
        assert np.max(np.abs(pos - po)) <= 4e-15 * (1 + np.max(np.abs(po)))
        assert np.max(np.abs(vel - vo)) <= 4e-15 * (1 + np.max(np.abs(vo)))


def _oracle_wake(sol, saved):
    w = sol.wake
    ow = WO.OracleWake(w.bordering_nodes, w.shedding_lines, w.adjacent_surfaces, sol.tol, 0.0, w.line_rows)
    ow.wake_positions[...] = saved["wake_positions"]
    ow.circulation[...] = saved["circulation"]
    ow.current_time = saved["current_time"]
    return ow


@pytest.mark.parametrize("frame,mode", [(RF.HeaveFrame(0.1, 0.9), "free"),
                                        (RF.ReferenceFrame((0.0, 0.0, 0.1), (0.0, 0.15, 0.0)), "reference")],
                         ids=["heave_free", "static_rotated"])
def test_postprocessing_matches_oracle(frame, mode):
    flow = FlowConditionsUniform(1.0, 0.0, 0.05)
    sol = make_wing_wake_solver(8, 4, 4.0, 1.0, frame, flow, 6, 0.125, mode=mode)
    res = run_solver(sol, TimeSettings(9, 0.125, 3))
    times = res.settings.time_settings.output_times
    assert times.size == 3 and len(res.wake_models) == 3
    pos0, el, _ = G.flat_wing(8, 4, 4.0, 1.0)
    topo = O.mesh_from_elements(pos0.shape[0], el)
    grid = np.ascontiguousarray(RNG.uniform([-0.5, -2.5, -0.5], [3.0, 2.5, 0.5], (400, 3)))
    vel = PP.compute_velocities(res, grid)
    vvar = PP.compute_velocities_variable(res, [grid[: 100 + 50 * i] for i in range(3)])
    psurf = PP.compute_surface_dynamic_pressure(res)
    pvar = PP.compute_dynamic_pressure_variable(res, [grid] * 3)
    forces = PP.circulatory_forces(res)
    vinf = np.array([1.0, 0.0, 0.05])
    for i, t in enumerate(times):
        pos = FO.positions_at_time(frame, pos0, t)
        mot = FO.velocity_at_time(frame, pos0, t)
        assert np.array_equal(res.geometry.positions_at_time(t), pos)
        assert np.array_equal(res.geometry.velocity_at_time(t), mot)
        circ = res.circulations[i]
        ow = _oracle_wake(sol, res.wake_models[i].save())
        assert _rel(vel[i], PO.velocity_at(topo, pos, circ, grid, vinf, 1e-6, ow)) < 1e-12
        g = grid[: 100 + 50 * i]
        assert _rel(vvar[i], PO.velocity_at(topo, pos, circ, g, vinf, 1e-6, ow)) < 1e-12
        assert _rel(psurf[i], PO.surface_pressure(topo, pos, mot, circ, vinf, 1.225, 1e-6, ow)) < 1e-12
        assert _rel(pvar[i], PO.pressure_at(topo, pos, circ, grid, vinf, 1.225, 1e-6, ow)) < 1e-12
        assert _rel(forces[i], PO.forces(topo, pos, mot, circ, vinf, 1e-6, ow)) < 1e-12


@pytest.mark.parametrize("mode", ["reference", "free"])
def test_rotating_free_stream_time_loop_matches_oracle(mode):
    """FlowConditionsRotating in the rhs and in the wake's convection (flow.get_velocity per
    node, python/pyvl/wake_models.py:124)."""
    flow = FlowConditionsRotating(0.5, -6.0, 0.0, 0.0, 0.0, -0.2)
    frame = RF.HeaveFrame(0.05, 0.7)
    dt, steps, R = 0.125, 9, 5
    sol = make_wing_wake_solver(8, 4, 4.0, 1.0, frame, flow, R, dt, mode=mode)
    pos0, el, _ = G.flat_wing(8, 4, 4.0, 1.0)
    topo = O.mesh_from_elements(pos0.shape[0], el)
    w = sol.wake
    ow = WO.OracleWake(w.bordering_nodes, w.shedding_lines, w.adjacent_surfaces, 1e-6, 0.0, R)
    field = WO.rotating_flow((0.5, -6.0, 0.0), (0.0, 0.0, -0.2))
    for k in range(steps):
        g = sol.step().numpy()
        go, _, _ = WO.step(k, dt, topo, pos0, frame, field, 1e-6, ow, free=(mode == "free"))
        assert _rel(g, go) < 1e-10, k
    wp, wc = w.state()
    assert _rel(wp, ow.wake_positions) < 1e-10 and _rel(wc, ow.circulation) < 1e-10


def test_rotor_frame_time_loop_matches_oracle():
    frame = RF.RotorReferenceFrame((0.0, 0.0, 0.0), (0.0, 0.05, 0.0), None, (0.0, 0.0, 0.3), 0.0)
    flow = FlowConditionsUniform(1.0, 0.0, 0.0)
    dt, steps, R = 0.125, 8, 5
    sol = make_wing_wake_solver(8, 4, 4.0, 1.0, frame, flow, R, dt)
    pos0, el, _ = G.flat_wing(8, 4, 4.0, 1.0)
    topo = O.mesh_from_elements(pos0.shape[0], el)
    w = sol.wake
    ow = WO.OracleWake(w.bordering_nodes, w.shedding_lines, w.adjacent_surfaces, 1e-6, 0.0, R)
    for k in range(steps):
        g = sol.step().numpy()
        go, _, _ = WO.step(k, dt, topo, pos0, frame, flow.vector, 1e-6, ow)
        assert _rel(g, go) < 1e-10, k


@pytest.mark.parametrize("fmt", ["json", "npz"])
def test_snapshot_resume_is_bit_identical(tmp_path, fmt):
    frame = RF.PitchHeaveFrame(np.radians(5.0), 0.05, 0.8, 0.25, 0.3, 1.1)
    flow = FlowConditionsUniform(1.0, 0.0, 0.0)

    def make():
        return make_wing_wake_solver(8, 4, 4.0, 1.0, frame, flow, 6, 0.125, mode="free")

    a = make()
    a.run(10)
    b = make()
    b.run(5)
    path = tmp_path / f"state.{fmt}"
    (fio.serialize_json if fmt == "json" else fio.serialize_npz)(b.snapshot().save(), path)
    state = SolverState.load((fio.deserialize_json if fmt == "json" else fio.deserialize_npz)(path))
    assert state.iteration == 5
    assert state.geometry.mesh.n_surfaces == 32
    assert type(state.geometry.rf) is RF.PitchHeaveFrame
    c = make()
    c.restore(state)
    c.run(5)
    assert np.array_equal(c.d_x.numpy(), a.d_x.numpy())
    pa, ca = a.wake.state()
    pc, cc = c.wake.state()
    assert np.array_equal(pa, pc) and np.array_equal(ca, cc)
    assert c.wake.step_count == a.wake.step_count
