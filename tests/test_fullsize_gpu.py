"""The benchmark workloads at their full sizes, checked against the oracle and LAPACK.

- 16,384-panel closed ellipsoid (4:1:0.5), 16 right-hand sides in the 10-degree cone about
  +x: FP64 LU pivots equal LAPACK's (scipy, the reference's solver,
  python/pyvl/solver.py:223-227), and the FP64 solve and both FP32-LU + FP64 refinement
  solves agree with LAPACK's FP64 solve of the same deflated matrix to 1e-10 (mean-free).
- 8,192-panel jittered UV sphere: the same for the FP64 LU (the pivoted tall-panel kernel).
- 64 x 16 wing with its 400-row wake full (25,600 wake panels): one step from the GPU's
  state, teacher forced, against the oracle's free-wake step and against the reference's
  own C core in the reference's convection semantics: circulation, wake positions and
  wake circulation to 1e-10 relative.
"""

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import oracle as O
from oracle import wake_oracle as WO
from pyvl_amd import DeviceArray
from pyvl_amd import geometry as G
from pyvl_amd import linalg as LA
from pyvl_amd.cvl import Mesh
from pyvl_amd.solver import FlowConditionsUniform, HeaveFrame, make_wing_wake_solver

pytestmark = pytest.mark.gpu


def _rel_cols(a, b):
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    return float(np.max(np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)))


def _mean_free(x):
    return x - x.mean(axis=0, keepdims=True)


def _closed_body_system(pos, el):
    m = Mesh(pos.shape[0], el)
    dp = DeviceArray.from_numpy(pos)
    nrm = m.surface_normal(dp)
    cps = m.surface_average_vec3(dp)
    n = m.n_surfaces
    a = DeviceArray((n, n))
    m.induction_matrix3(1e-6, dp, cps, nrm, a)
    LA.deflate(a, 0, n, LA.deflation_constant(np.diag(a.numpy())))
    return a, nrm.numpy()


def test_c4_16k_closed_body_solves_match_lapack():
    pos, el = G.ellipsoid(128, 128, axes=(4.0, 1.0, 0.5), sigma_rel=1e-3, seed=0)
    a, nrm = _closed_body_system(pos, el)
    n = a.shape[0]
    assert n == 16384
    dirs = np.stack([G.cone_directions(1, 10.0, seed)[0] for seed in range(1, 17)])
    b = np.ascontiguousarray(-(nrm @ dirs.T))
    db = DeviceArray.from_numpy(b)
    ad = a.numpy()
    lu_ref, piv_ref = sla.lu_factor(ad, check_finite=False)
    x_ref = _mean_free(sla.lu_solve((lu_ref, piv_ref), b, check_finite=False))
    del lu_ref

    lu64 = LA.LU(n, LA.FP64).factor(a)
    assert np.array_equal(lu64.pivots(), piv_ref)
    x64 = DeviceArray((n, 16))
    import ctypes as C
    from pyvl_amd import _lib

    _lib.call("pvl_memcpy_d2d", x64.cptr, db.cptr, C.c_size_t(db.nbytes))
    lu64.solve(x64)
    assert _rel_cols(_mean_free(x64.numpy()), x_ref) < 1e-10
    del lu64

    for pivoting in (True, False):
        lu32 = LA.LU(n, LA.FP32, pivoting=pivoting).factor(a)
        x, iters, _, res = LA.solve_refine(a, db, lu32, tol=1e-12)
        assert np.all(res <= 1e-12) and iters <= 8, (pivoting, iters, res)
        assert _rel_cols(_mean_free(x.numpy()), x_ref) < 1e-10, pivoting
        del lu32


def test_c2_sphere_8k_lu_matches_lapack():
    pos, el = G.uv_sphere(128, 64, sigma=1e-3, seed=0)
    a, nrm = _closed_body_system(pos, el)
    n = a.shape[0]
    assert n == 8192
    b = -(nrm @ np.array([1.0, 0.0, 0.0]))
    ad = a.numpy()
    lu_ref, piv_ref = sla.lu_factor(ad, check_finite=False)
    h = LA.LU(n, LA.FP64).factor(a)
    assert np.array_equal(h.pivots(), piv_ref)
    x = h.solve(b)
    x_ref = sla.lu_solve((lu_ref, piv_ref), b, check_finite=False)
    x -= x.mean()
    x_ref -= x_ref.mean()
    assert float(np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref)) < 1e-10


C3 = dict(n_span=64, n_chord=16, span=8.0, chord=1.0, h0=0.1, omega=0.6, dt=1.0 / 16.0, rows=400, tol=1e-6)


def _c3_solver(mode):
    fc = FlowConditionsUniform(1.0, 0.0, 0.0)
    frame = HeaveFrame(C3["h0"] * C3["chord"], C3["omega"])
    sol = make_wing_wake_solver(C3["n_span"], C3["n_chord"], C3["span"], C3["chord"], frame, fc, C3["rows"],
                                C3["dt"], vortex_limit=C3["tol"], mode=mode)
    sol.run(C3["rows"] + 2)
    return sol


def _oracle_from(sol):
    el_counts, flat = sol.mesh.to_element_connectivity()
    offs = np.concatenate([[0], np.cumsum(el_counts)]).astype(np.int64)
    elements = [flat[offs[i]:offs[i + 1]].tolist() for i in range(len(el_counts))]
    bound = O.mesh_from_elements(sol.p0.shape[0], elements)
    w = sol.wake
    ow = WO.OracleWake(w.bordering_nodes, w.shedding_lines, w.adjacent_surfaces, C3["tol"], 0.0, C3["rows"])
    wp, wc = w.state()
    ow.wake_positions[...] = wp
    ow.circulation[...] = wc
    ow.step_count = w.step_count
    ow.current_time = w.current_time
    return bound, ow


@pytest.mark.parametrize("mode", ["free", "reference"])
def test_c3_full_wake_step_matches_oracle(mode):
    if mode == "reference" and O.load_ref() is None:
        pytest.skip("oracle/_ref (the reference's own C core) is not built")
    sol = _c3_solver(mode)
    bound, ow = _oracle_from(sol)
    it = sol.iteration
    vinf = np.array([1.0, 0.0, 0.0])
    if mode == "free":
        g_o, _, _ = WO.step(it, sol.dt, bound, sol.p0, sol.frame, vinf, C3["tol"], ow, free=True)
    else:
        g_o = WO.reference_core_step(it, sol.dt, bound, sol.p0, sol.frame, vinf, C3["tol"], ow)
    g = sol.step().numpy()
    wp, wc = sol.wake.state()
    assert _rel_cols(g[:, None], g_o[:, None]) < 1e-10
    assert float(np.max(np.abs(wp - ow.wake_positions)) / np.max(np.abs(ow.wake_positions))) < 1e-10
    assert float(np.max(np.abs(wc - ow.circulation)) / np.max(np.abs(ow.circulation))) < 1e-10
    assert sol.wake.step_count == ow.step_count


def test_c5_full_run_gmres_matches_fp64_lu_and_oracle():
    """The learned-start workload at its full size: tapered wing 64 x 32 = 2,048 panels,
    pitch + heave (kinematics seed 7), free wake of 300 rows, 300 steps.  GMRES-IR from the
    minimal-residual projection start (the default iterative start) against the FP64 LU run
    at every step (<= 1e-10, no LU fallback), and at two sampled steps both against the
    oracle's FP64 solve (scipy LU) of the step's system, assembled on the CPU from the wake
    state the GPU run had at that step (<= 1e-10)."""
    from pyvl_amd import learned as L
    from pyvl_amd.solver import make_c5_solver
    from oracle import frames_oracle as FO

    steps, rows = 300, 300
    judged = {}

    def capture(sol, it):
        if it in (150, steps - 1):
            wp, wc = sol.wake.state()
            judged[it] = (sol.d_pos.numpy(), wp, wc)

    ref = L.run_c5(7, steps, solve="lu64", rows=rows)
    run = L.run_c5(7, steps, solve="gmres", warm="proj", rows=rows, pre_solve=capture)
    assert run["fallbacks"] == 0
    err = np.max(np.abs(run["gamma"] - ref["gamma"]), axis=1) / np.max(np.abs(ref["gamma"]), axis=1)
    assert np.max(err) < 1e-10
    sol = make_c5_solver(7, rows=rows)
    pos0, elements, _ = G.flat_wing(64, 32, 8.0, 1.0, taper=0.5)
    topo = O.mesh_from_elements(pos0.shape[0], elements)
    w = sol.wake
    for it, (pos, wp, wc) in judged.items():
        cps, nrm = O.surface_center(topo, pos), O.surface_normal(topo, pos)
        vel = FO.kinematics(sol.frame, sol.p0, it * sol.dt)[1]
        b = np.vecdot(nrm, -(np.array([1.0, 0.0, 0.0]) - O.surface_center(topo, vel)), axis=1)
        a = O.induction_matrix3(topo, pos, cps, nrm, 1e-6)
        ow = WO.OracleWake(w.bordering_nodes, w.shedding_lines, w.adjacent_surfaces, 1e-6, 0.0, rows)
        ow.wake_positions[...] = wp
        ow.circulation[...] = wc
        ow.apply_corrections(cps, nrm, a, b)
        g_or = sla.lu_solve(sla.lu_factor(a), b)
        for g in (run["gamma"][it], ref["gamma"][it]):
            assert float(np.linalg.norm(g - g_or) / np.linalg.norm(g_or)) < 1e-10, it
