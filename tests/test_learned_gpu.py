"""Learned warm start on the GPU: the MLP kernels against the CPU oracle, GMRES refinement
against the FP64 LU, and what the trained network gains over the method without it.

Tolerances (BF16 operands, FP32 accumulation): network output within 1e-2 relative
(Frobenius) of the oracle that rounds at the same points, gradients within 2e-2 per layer
of it and within 6e-2 of the exact float64 gradient; Adam within 1e-5 * lr of its formula
applied to the device gradient.  Solves: 1e-10 relative to the FP64 LU.
"""

import ctypes as C

import numpy as np
import pytest

from oracle import frames_oracle as FO
from oracle import mlp_oracle as MO
from pyvl_amd import DeviceArray, _lib
from pyvl_amd import geometry as G
from pyvl_amd import learned as L
from pyvl_amd import linalg as LA
from pyvl_amd.cvl import Mesh

pytestmark = pytest.mark.gpu

D_IN = len(L.FEATURES)
DIMS = (D_IN, 256, 256, 256, 1)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _net(seed=1):
    net = L.MLP(D_IN, 256, 4, seed=seed)
    p = net.get_params().astype(np.float64)
    # non-zero biases so the bias paths are exercised
    rng = np.random.default_rng(seed)
    for w, b in net.unpack(p):
        b[:] = rng.normal(0.0, 0.1, b.size)
    net.set_params(p.astype(np.float32))
    return net, net.get_params().astype(np.float64)


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, D_IN)).astype(np.float32)
    t = (np.sin(X[:, 0]) + X[:, 1] * X[:, 2] - 0.5 * X[:, 3]).astype(np.float32)
    return X, t


@pytest.mark.parametrize("batch", [1, 63, 300, 2048])
def test_forward_matches_oracle(batch):
    net, p = _net()
    X, _ = _data(batch, batch)
    y = net.forward(DeviceArray.from_numpy(X)).numpy()
    y_ref, _, _ = MO.forward(p, DIMS, X, bf16=True)
    assert _rel(y, y_ref) < 1e-2
    y_exact, _, _ = MO.forward(p, DIMS, X)
    assert _rel(y, y_exact) < 3e-2


@pytest.mark.parametrize("batch", [64, 777])
def test_gradient_matches_oracle(batch):
    net, p = _net(3)
    X, t = _data(batch, 10 + batch)
    g, loss = net.gradient(DeviceArray.from_numpy(X), DeviceArray.from_numpy(t))
    loss_ref, g_ref = MO.loss_and_gradient(p, DIMS, X, t, bf16=True)
    _, g_exact = MO.loss_and_gradient(p, DIMS, X, t)
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref
    off = 0
    for l in range(len(DIMS) - 1):
        nw = DIMS[l] * DIMS[l + 1]
        for sl in (slice(off, off + nw), slice(off + nw, off + nw + DIMS[l + 1])):
            assert _rel(g[sl], g_ref[sl]) < 2e-2, (l, sl)
            assert _rel(g[sl], g_exact[sl]) < 6e-2, (l, sl)
        off += nw + DIMS[l + 1]
    assert off == g.size


def test_adam_step_matches_oracle():
    net, p = _net(5)
    X, t = _data(512, 5)
    dX, dt = DeviceArray.from_numpy(X), DeviceArray.from_numpy(t)
    g, _ = net.gradient(dX, dt)
    lr, wd = 1e-3, 1e-4
    net.train_step(dX, dt, lr, wd=wd)
    p1 = net.get_params()
    p_ref, _, _ = MO.adam(p, g.astype(np.float64), np.zeros_like(p), np.zeros_like(p), 1, lr, wd=wd)
    assert np.max(np.abs(p1 - p_ref)) <= 1e-5 * lr + 1e-7 * np.max(np.abs(p))


def test_training_reduces_loss():
    net = L.MLP(D_IN, 256, 4, seed=11)
    X, t = _data(8192, 11)
    dX, dt = DeviceArray.from_numpy(X), DeviceArray.from_numpy(t)
    first = net.train_step(dX, dt, 1e-3, want_loss=True)
    for _ in range(150):
        last = net.train_step(dX, dt, 1e-3, want_loss=True)
    assert last < 0.1 * first


def test_bad_batch_is_an_error():
    net = L.MLP(14, 64, 3)
    x = DeviceArray((1, 14), np.float32)
    with pytest.raises(ValueError):
        net.gradient(x.view(0, (0, 14)), DeviceArray((1,), np.float32).view(0, (0,)))


# ---------------------------------------------------------------------------------------
def _wing_system(n_span=32, n_chord=8):
    pos, elements, _ = G.flat_wing(n_span, n_chord, 8.0, 1.0, taper=0.5)
    m = Mesh(pos.shape[0], elements)
    aoa = np.radians(5.0)
    p = pos.copy()
    p[:, 2] = -np.sin(aoa) * p[:, 0]
    p[:, 0] = np.cos(aoa) * p[:, 0]
    cps = m.surface_average_vec3(p)
    nrm = m.surface_normal(p)
    A = m.induction_matrix3(1e-6, p, cps, nrm)
    b = -nrm @ np.array([1.0, 0.0, 0.0])
    return A, b


def _gmres(A, b, x0=None, **kw):
    n = b.size
    dA = DeviceArray.from_numpy(A)
    dA32 = DeviceArray((n, n), np.float32)
    _lib.call("pvl_convert_f64_f32", dA.cptr, dA32.cptr, C.c_size_t(n * n))
    dx = DeviceArray.from_numpy(np.zeros(n) if x0 is None else x0)
    ok, outer, steps, rel0, rel1 = L.gmres_refine(dA, dA32, DeviceArray.from_numpy(b), dx, **kw)
    x = dx.numpy()
    # the reported residual is the true one
    if np.any(b):
        r = np.max(np.abs(b - A @ x)) / np.max(np.abs(b))
        assert rel1 == pytest.approx(r, rel=1e-2, abs=1e-15)
    return ok, outer, steps, rel0, rel1, x


def test_gmres_wing_matches_lu():
    A, b = _wing_system()
    x_ref = LA.LU(b.size).factor(A).solve(b)
    ok, outer, steps, rel0, rel1, x = _gmres(A, b)
    assert ok and rel0 == 1.0 and rel1 <= 1e-12
    assert _rel(x, x_ref) < 1e-10
    # started near the answer: fewer corrections and GMRES steps than from zero
    x_near = x_ref * (1.0 + 1e-6 * np.random.default_rng(0).normal(size=b.size))
    ok2, outer2, steps2, rel02, _, x2 = _gmres(A, b, x_near)
    assert ok2 and rel02 < 1e-5 and outer2 < outer and steps2 < steps
    assert _rel(x2, x_ref) < 1e-10
    # started at the answer: no correction at all
    ok3, outer3, steps3, _, _, _ = _gmres(A, b, x)
    assert ok3 and outer3 == 0 and steps3 == 0


def test_gmres_time_loop_default_start_is_projection():
    """UnsteadySolver(solve="gmres") starts each solve from the projection on its last three
    circulations by default: the same circulations as the FP64 LU loop to 1e-10, and from the
    third step on a start residual far below the zero start's 1."""
    from pyvl_amd.solver import FlowConditionsUniform, PitchHeaveFrame, make_wing_wake_solver

    frame = PitchHeaveFrame(np.radians(5.0), 0.05, 0.8, 0.25, 0.3, 1.1)
    flow = FlowConditionsUniform(1.0, 0.0, 0.0)
    a = make_wing_wake_solver(16, 4, 4.0, 1.0, frame, flow, 10, 0.0625, mode="free", solve="gmres")
    b = make_wing_wake_solver(16, 4, 4.0, 1.0, frame, flow, 10, 0.0625, mode="free")
    assert isinstance(a.warm_start, L.ProjectionStart)
    for k in range(12):
        ga, gb = a.step().numpy(), b.step().numpy()
        assert _rel(ga, gb) < 1e-10, k
    rel0 = [st[2] for st in a.gmres_stats]
    assert rel0[0] == 1.0 and max(rel0[3:]) < 1e-3
    assert a.fallbacks == 0


@pytest.mark.parametrize("n", [1, 100, 1000, 1537])
def test_gmres_random_matches_lu(n):
    rng = np.random.default_rng(n)
    A = rng.normal(size=(n, n)) / np.sqrt(n) + 2.0 * np.eye(n) * (1.0 + rng.uniform(size=n))
    b = rng.normal(size=n)
    x_ref = LA.LU(n).factor(A).solve(b)
    ok, _, _, _, _, x = _gmres(A, b)
    assert ok
    assert _rel(x, x_ref) < 1e-10


def test_gmres_zero_rhs_and_non_convergence():
    """b = 0 gives x = 0; a budget too small to converge returns not-converged, which the
    solver turns into the LU fallback; a NaN start is reported, never accepted."""
    A, b = _wing_system()
    ok, outer, _, _, _, x = _gmres(A, np.zeros_like(b), np.ones_like(b))
    assert ok and outer == 0 and not np.any(x)
    ok, outer, steps, _, rel1, _ = _gmres(A, b, m=2, max_outer=1)
    assert not ok and outer == 1 and rel1 > 1e-12
    n = b.size
    dA = DeviceArray.from_numpy(A)
    dA32 = DeviceArray((n, n), np.float32)
    _lib.call("pvl_convert_f64_f32", dA.cptr, dA32.cptr, C.c_size_t(n * n))
    x0 = np.zeros(n)
    x0[3] = np.nan
    ok, outer, _, _, _ = L.gmres_refine(dA, dA32, DeviceArray.from_numpy(b), DeviceArray.from_numpy(x0))
    assert not ok and outer == 0


def test_learned_warm_start_gains():
    """Reduced pitch + heave workload (2,048 panels, 40-row free wake): a network trained on
    eight other kinematics seeds (150 steps each) starts GMRES closer to the solution than the
    linear extrapolation it corrects -- the method without it -- (which beats the previous
    step, which beats zero) and needs fewer GMRES steps than the zero start; every step's
    circulation stays
    within 1e-10 of the FP64 LU run, and at sampled steps within 1e-10 of the oracle's FP64
    solve (scipy LU) of the same step's system, assembled on the CPU from the wake state the
    GPU run had at that step.  The minimal-residual start over the last three circulations
    is reported beside it (the network adds nothing measurable on top of that one)."""
    import scipy.linalg as sla

    from oracle import oracle as O
    from oracle.wake_oracle import OracleWake

    rows, steps = 40, 40
    X, y = L.generate_training_data(n_steps=8 * 150, seeds=tuple(range(100, 108)), rows=rows)
    norm = L.Normaliser(X, y)
    net = L.MLP(X.shape[1], 256, 4, seed=42)
    losses = L.train(net, X, y, norm, epochs=12, batch=8192, lr=2e-3)
    print(f"training losses {losses}")
    assert losses[-1] < 0.2 * losses[0]
    judged = {}

    def capture(sol, it):
        if it in (5, steps - 1):
            wp, wc = sol.wake.state()
            judged[it] = (sol.d_pos.numpy(), wp, wc)

    ref = L.run_c5(7, steps, solve="lu64", rows=rows)
    runs = {w: L.run_c5(7, steps, solve="gmres", warm=w, mlp=net, norm=norm, rows=rows,
                        pre_solve=capture if w == "mlp" else None)
            for w in ("none", "prev", "extrap", "mlp", "proj", "proj_mlp")}
    for r in runs.values():
        assert r["fallbacks"] == 0
        err = np.max(np.abs(r["gamma"] - ref["gamma"]), axis=1) / np.max(np.abs(ref["gamma"]), axis=1)
        assert np.max(err) < 1e-10
    # from step 2 on every start has its full history
    steps_ = {w: np.mean([s[1] for s in r["gmres"][2:]]) for w, r in runs.items()}
    res0 = {w: np.median([s[2] for s in r["gmres"][2:]]) for w, r in runs.items()}
    print(f"GMRES steps {steps_}, median start residual {res0}")
    assert res0["extrap"] < res0["prev"] < res0["none"]
    assert res0["mlp"] < res0["extrap"], res0
    # fewer GMRES steps than the cold start; not fewer than the extrapolation: each
    # correction runs to the FP32 floor, and the network's remainder is rougher
    assert steps_["mlp"] < steps_["none"], steps_
    warm = runs["mlp"]
    # the oracle judges the learned run's circulation at the captured steps
    from pyvl_amd.solver import make_c5_solver

    sol = make_c5_solver(7, rows=rows)
    pos0, elements, _ = G.flat_wing(64, 32, 8.0, 1.0, taper=0.5)
    topo = O.mesh_from_elements(pos0.shape[0], elements)
    w = sol.wake
    for it, (pos, wp, wc) in judged.items():
        cps, nrm = O.surface_center(topo, pos), O.surface_normal(topo, pos)
        vel = FO.kinematics(sol.frame, sol.p0, it * sol.dt)[1]
        b = np.vecdot(nrm, -(np.array([1.0, 0.0, 0.0]) - O.surface_center(topo, vel)), axis=1)
        a = O.induction_matrix3(topo, pos, cps, nrm, 1e-6)
        ow = OracleWake(w.bordering_nodes, w.shedding_lines, w.adjacent_surfaces, 1e-6, 0.0, rows)
        ow.wake_positions[...] = wp
        ow.circulation[...] = wc
        if it > 0:
            ow.apply_corrections(cps, nrm, a, b)
        g_or = sla.lu_solve(sla.lu_factor(a), b)
        assert _rel(warm["gamma"][it], g_or) < 1e-10, it
