#!/usr/bin/env python3
"""Benchmark of the line-vortex wake step and the 16k-panel solve on one MI355X.

Headline (one JSON line on stdout): wake steps/s of the unsteady free-wake workload --
rectangular flat wing, 64 x 16 = 1,024 bound panels (span 8, chord 1), heave 0.1 chord at
reduced frequency k = 0.3 (omega = 0.6), dt = c/(16 U), one wake row shed per step into a
400-row wake (25,600 wake panels; 26,065 distinct wake nodes, 51,664 distinct segments),
free stream 1 m/s along x, vortex cut-off 1e-6, FP64 throughout.  A step is the reference's
run_solver iteration (python/pyvl/solver.py:175-239): kinematics, rhs, system matrix,
wake corrections, LU solve and wake update, with free-wake convection (free stream + bound
+ wake self-induction at every wake node).  The wake is grown to its full 400 rows first;
`--warmup` further full-wake steps are untimed and `--steps` are timed.

Also reported: the same workload with FP32 induction angles, the reference's own
convection semantics, and the 16k-panel closed-body solve (ellipsoid 4:1:0.5, 16 right-hand
sides): plain FP64 LU against FP32 LU on MFMA + FP64 iterative refinement, with the error
of both against LAPACK FP64 (scipy, as the reference calls it) on the host.

CPU baseline: the oracle's C restatement (OpenMP, all host cores) of the same free-wake
step, on a bounded sample (one full-wake step started from the GPU's wake state).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and MFMA)
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 (vector and MFMA)
# HBM bytes per launch of the full-wake self-induction kernel: rocprofv3 FETCH_SIZE and
# WRITE_SIZE in two runs of this script (--copy-calibration), each scaled by what the same
# counter reads for the script's own 1 GiB copy (FETCH x 2.0, WRITE x 1.0); the commands and
# numbers are in profiles/wake_self_velocity_c3_fp64_traffic.md.
SELF_VELOCITY_HBM_BYTES = 71.82e6
PEAK_HBM_GBS = 8000.0
FLOP_PER_PAIR = 87.0  # induction pair (src/core/flow_solver.c:114-135, atan2 counted as 20)
FLOP_WEIGHT = 6.0  # Gamma weighting of a point-velocity sum

C3 = dict(n_span=64, n_chord=16, span=8.0, chord=1.0, h0=0.1, omega=0.6, dt=1.0 / 16.0, rows=400, tol=1e-6)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def prof_get(lib, name: str):
    ms, cnt, work = C.c_double(), C.c_long(), C.c_double()
    lib.pvl_prof_query(name.encode(), C.byref(ms), C.byref(cnt))
    lib.pvl_prof_work(name.encode(), C.byref(work))
    return ms.value, cnt.value, work.value


def run_c3(mode: str, prec: int, steps: int, warmup: int, profile: bool):
    from pyvl_amd import _lib
    from pyvl_amd.solver import FlowConditionsUniform, HeaveFrame, make_wing_wake_solver

    lib = _lib.load()
    fc = FlowConditionsUniform(1.0, 0.0, 0.0)
    frame = HeaveFrame(C3["h0"] * C3["chord"], C3["omega"])
    sol = make_wing_wake_solver(C3["n_span"], C3["n_chord"], C3["span"], C3["chord"], frame, fc, C3["rows"],
                                C3["dt"], vortex_limit=C3["tol"], mode=mode, induction_precision=prec)
    t0 = time.perf_counter()
    grow = sol.run(C3["rows"])
    grow_s = time.perf_counter() - t0
    log(f"{mode}/{'fp64' if prec == 0 else 'fp32'}: wake grown to {C3['rows']} rows in {grow_s:.2f} s "
        f"(last 50 growth steps {np.mean(grow[-50:]) * 1e3:.2f} ms/step)")
    sol.run(warmup)
    if profile:
        # inside the timed region only the dominant kernel carries HIP events
        lib.pvl_prof_reset()
        lib.pvl_prof_filter(b"wake_self_velocity")
        lib.pvl_prof_enable(1)
    _lib.call("pvl_synchronize")
    t0 = time.perf_counter()
    for _ in range(steps):
        sol.step()
    _lib.call("pvl_synchronize")
    timed = time.perf_counter() - t0
    prof, breakdown = {}, {}
    if profile:
        prof["wake_self_velocity"] = prof_get(lib, "wake_self_velocity")
        # per-kernel breakdown from further (untimed) full-wake steps with every scope timed
        lib.pvl_prof_filter(None)
        lib.pvl_prof_reset()
        n_b = min(steps, 10)
        for _ in range(n_b):
            sol.step()
        _lib.call("pvl_synchronize")
        for name in ("wake_self_velocity", "wake_bound_velocity", "wake_explicit", "induction_matrix3",
                     "lu_panel_f64", "lu_laswp_f64", "lu_trsm_f64", "lu_gemm_f64", "lu_trsv_f64"):
            ms, cnt, _ = prof_get(lib, name)
            breakdown[name] = {"ms_per_step": round(ms / n_b, 4), "launches_per_step": cnt / n_b}
        lib.pvl_prof_enable(0)
    log(f"{mode}/{'fp64' if prec == 0 else 'fp32'}: {steps} full-wake steps in {timed:.3f} s "
        f"-> {steps / timed:.1f} steps/s")
    return dict(sol=sol, timed_s=timed, steps_per_s=steps / timed, ms_per_step=1e3 * timed / steps,
                grow_400_s=grow_s, prof=prof, breakdown=breakdown)


def roofline_from(prof) -> dict:
    ms, cnt, pairs = prof["wake_self_velocity"]
    flops = pairs * (FLOP_PER_PAIR + FLOP_WEIGHT)
    avg_ms = ms / max(cnt, 1)
    achieved = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    return {
        "bound": "valu_fp64",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
        "traffic": SELF_VELOCITY_HBM_BYTES,
        "kernel": "chain_velocity_kernel<FP64> (free-wake self-induction over segment chains)",
        "note": "FP64 vector-ALU bound (peak: MI355X FP64 vector rate, 78.6 TFLOP/s, equal to its FP64 MFMA rate); "
                f"algorithmic {FLOP_PER_PAIR + FLOP_WEIGHT:.0f} FLOP per (segment, point) pair, "
                f"{pairs / max(cnt, 1):.4g} pairs per launch, avg {avg_ms:.3f} ms over {cnt} launches "
                "(HIP events on the side stream the kernel is launched on, around this kernel only, over "
                "the timed region; it runs beside the bound solve, capped at 5 workgroups per CU)",
    }


def _oracle_from(sol):
    """The oracle's bound mesh and reference-layout wake, loaded with the GPU solver's state."""
    from oracle import oracle as O
    from oracle.wake_oracle import OracleWake

    el_counts, flat = sol.mesh.to_element_connectivity()
    offs = np.concatenate([[0], np.cumsum(el_counts)]).astype(np.int64)
    elements = [flat[offs[i]:offs[i + 1]].tolist() for i in range(len(el_counts))]
    bound = O.mesh_from_elements(sol.p0.shape[0], elements)
    w = sol.wake
    ow = OracleWake(w.bordering_nodes, w.shedding_lines, w.adjacent_surfaces, C3["tol"], 0.0, C3["rows"])
    wp, wc = w.state()
    ow.wake_positions[...] = wp
    ow.circulation[...] = wc
    ow.step_count = w.step_count
    ow.current_time = w.current_time
    return bound, ow


def _step_errors(sol, gammas_cpu, ow, n_steps) -> dict:
    """The GPU solver's next n_steps from the same state, against the host's."""
    err_g = 0.0
    for k in range(n_steps):
        g = sol.step().numpy()
        err_g = max(err_g, float(np.linalg.norm(g - gammas_cpu[k]) / np.linalg.norm(gammas_cpu[k])))
    wp, wc = sol.wake.state()
    return {
        "rel_err_circulation": err_g,
        "rel_err_wake_positions": float(np.max(np.abs(wp - ow.wake_positions)) / np.max(np.abs(ow.wake_positions))),
        "rel_err_wake_circulation": float(np.max(np.abs(wc - ow.circulation)) / np.max(np.abs(ow.circulation))),
    }


def cpu_free_wake_step(gpu_run) -> dict:
    """One full-wake free-wake step on the host, oracle C restatement (reference layout),
    from the GPU's state; the GPU's own next step is then compared with it."""
    from oracle import wake_oracle as WO

    sol = gpu_run["sol"]
    bound, ow = _oracle_from(sol)
    it = sol.iteration
    vinf = np.array([1.0, 0.0, 0.0])
    t0 = time.perf_counter()
    gamma, _, _ = WO.step(it, sol.dt, bound, sol.p0, sol.frame, vinf, C3["tol"], ow, free=True)
    elapsed = time.perf_counter() - t0
    errs = _step_errors(sol, [gamma], ow, 1)
    return {
        "value": round(1.0 / elapsed, 4),
        "unit": "steps/s",
        "cores": os.cpu_count(),
        "kind": "port",
        "sample": f"1 full-wake free-wake step (step {it + 1}, 25,600 wake panels, reference layout: "
                  f"{ow.wake_positions[:, :, :-1].size // 3} wake nodes x {ow.wake_mesh.n_lines} wake lines) from the "
                  f"GPU's wake state; oracle C restatement, gcc -O3 no FMA, OpenMP {os.cpu_count()} threads; "
                  f"LU = scipy LAPACK; {elapsed:.2f} s",
        "gpu_vs_cpu": errs,
    }


def cpu_reference_step(gpu_ref_run, n_steps: int) -> dict:
    """Reference-semantics steps on the host with the reference's own C core (oracle/_ref),
    from the GPU's state; the GPU's own next steps are then compared with them."""
    from oracle import oracle as O
    from oracle import wake_oracle as WO

    if O.load_ref() is None:
        return {"skipped": "oracle/_ref not built"}
    sol = gpu_ref_run["sol"]
    bound, ow = _oracle_from(sol)
    vinf = np.array([1.0, 0.0, 0.0])
    it0 = sol.iteration
    gammas = []
    t0 = time.perf_counter()
    for k in range(n_steps):
        gammas.append(WO.reference_core_step(it0 + k, sol.dt, bound, sol.p0, sol.frame, vinf, C3["tol"], ow))
    elapsed = time.perf_counter() - t0
    errs = _step_errors(sol, gammas, ow, n_steps)
    return {
        "value": round(n_steps / elapsed, 4),
        "unit": "steps/s",
        "cores": os.cpu_count(),
        "kind": "reference",
        "sample": f"{n_steps} full-wake steps with the reference's convection (bound induction only), the "
                  "reference's own src/core C (compute_line_induction + line_induction_to_*_surface_induction "
                  "with its full line buffer, built unmodified into oracle/_ref), its Python step order restated "
                  f"in numpy, scipy LAPACK LU; {elapsed:.2f} s",
        "gpu_vs_cpu": errs,
    }


def fp32_vs_fp64_step(sol32, sol64) -> dict:
    """Error of the FP32-angle induction: the FP64 solver is loaded with the FP32 run's wake
    state and both take the same next step (the FP64 step is itself within 1e-12 of the
    oracle's); relative differences of circulation, wake positions and wake circulation."""
    import ctypes as C

    from pyvl_amd import _lib

    d = sol32.wake.save()
    pos = np.ascontiguousarray(d["wake_positions"], np.float64)
    circ = np.ascontiguousarray(d["circulation"], np.float64)
    _lib.call("pvl_wake_set_state", sol64.wake._h, pos.ctypes.data_as(C.c_void_p), circ.ctypes.data_as(C.c_void_p),
              C.c_uint(int(d["step_count"])), C.c_double(float(d["current_time"])))
    sol64.iteration = sol32.iteration
    g64 = sol64.step().numpy()
    g32 = sol32.step().numpy()
    p64, c64 = sol64.wake.state()
    p32, c32 = sol32.wake.state()
    return {
        "rel_err_circulation": float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
        "rel_err_wake_positions": float(np.max(np.abs(p32 - p64)) / np.max(np.abs(p64))),
        "rel_err_wake_circulation": float(np.max(np.abs(c32 - c64)) / np.max(np.abs(c64))),
    }


def run_c4(do_cpu: bool) -> dict:
    """16k-panel closed-body solve: ellipsoid 4:1:0.5, 128 x 128 (16,384 panels), 16 RHS."""
    from pyvl_amd import DeviceArray, _lib
    from pyvl_amd import geometry as G
    from pyvl_amd import linalg as LA
    from pyvl_amd.cvl import Mesh

    lib = _lib.load()
    pos, el = G.ellipsoid(128, 128, axes=(4.0, 1.0, 0.5), sigma_rel=1e-3, seed=0)
    mesh = Mesh(pos.shape[0], el)
    n = mesh.n_surfaces
    dp = DeviceArray.from_numpy(pos)
    nrm = mesh.surface_normal(dp)
    cps = mesh.surface_average_vec3(dp)
    A = DeviceArray((n, n))
    mesh.induction_matrix3(1e-6, dp, cps, nrm, A)  # warm
    _lib.call("pvl_synchronize")
    lib.pvl_prof_reset()
    lib.pvl_prof_enable(1)
    mesh.induction_matrix3(1e-6, dp, cps, nrm, A)
    _lib.call("pvl_synchronize")
    asm_ms, _, asm_pairs = prof_get(lib, "induction_matrix3")
    lib.pvl_prof_enable(0)
    diag = np.empty(n)
    # deflation constant c = mean |diag A| / n (host copy of the diagonal)
    Ah = A.numpy()
    c = LA.deflation_constant(np.diag(Ah))
    LA.deflate(A, 0, n, c)
    dirs = np.stack([G.cone_directions(1, 10.0, seed)[0] for seed in range(1, 17)])
    nh = nrm.numpy()
    B = np.ascontiguousarray(-(nh @ dirs.T))  # b_q = n . -(V_q)
    dB = DeviceArray.from_numpy(B)

    def mean_free(x):
        return x - x.mean(axis=0, keepdims=True)

    # plain FP64
    lu64 = LA.LU(n, LA.FP64)
    X64 = DeviceArray((n, 16))
    lu64.factor(A)
    _lib.call("pvl_synchronize")
    t0 = time.perf_counter()
    lu64.factor(A)
    _lib.call("pvl_memcpy_d2d", X64.cptr, dB.cptr, C.c_size_t(dB.nbytes))
    lu64.solve(X64)
    _lib.call("pvl_synchronize")
    t64 = time.perf_counter() - t0
    x64 = mean_free(X64.numpy())
    log(f"C4 FP64 LU+solve (16 RHS): {t64 * 1e3:.1f} ms")
    # FP32 LU (MFMA) + FP64 refinement
    lu32 = LA.LU(n, LA.FP32)
    lu32.factor(A)
    X32 = DeviceArray((n, 16))
    _lib.call("pvl_synchronize")
    t0 = time.perf_counter()
    lu32.factor(A)
    _, iters, _, relres = LA.solve_refine(A, dB, lu32, tol=1e-12, out=X32)
    _lib.call("pvl_synchronize")
    t32 = time.perf_counter() - t0
    x32 = mean_free(X32.numpy())
    log(f"C4 FP32 LU + FP64 IR (16 RHS): {t32 * 1e3:.1f} ms, {iters} refinement steps")
    # FP32 LU without pivoting (no per-column row search) + FP64 refinement
    lu32n = LA.LU(n, LA.FP32, pivoting=False)
    lu32n.factor(A)
    X32n = DeviceArray((n, 16))
    _lib.call("pvl_synchronize")
    t0 = time.perf_counter()
    lu32n.factor(A)
    _, iters_n, _, relres_n = LA.solve_refine(A, dB, lu32n, tol=1e-12, out=X32n)
    _lib.call("pvl_synchronize")
    t32n = time.perf_counter() - t0
    x32n = mean_free(X32n.numpy())
    log(f"C4 FP32 LU without pivoting + FP64 IR (16 RHS): {t32n * 1e3:.1f} ms, {iters_n} refinement steps")
    # per-kernel split of one more factorisation + solve of each kind (HIP events on every
    # launch, untimed above); roofline of the whole solve and of its MFMA trailing update
    kinds = {"fp64": (lu64, X64, 8, PEAK_FP64_TFLOPS, t64), "fp32_partial_pivoting_ir": (lu32, X32, 4, PEAK_FP32_TFLOPS, t32),
             "fp32_no_pivoting_ir": (lu32n, X32n, 4, PEAK_FP32_TFLOPS, t32n)}
    roof = {}
    nrhs = 16
    alg_flops = 2.0 / 3.0 * n ** 3 + 2.0 * n * n * nrhs
    for name, (lu, X, es, peak, tw) in kinds.items():
        sfx = "f64" if es == 8 else "f32"
        lib.pvl_prof_reset()
        lib.pvl_prof_enable(1)
        if es == 8:
            lu.factor(A)
            _lib.call("pvl_memcpy_d2d", X.cptr, dB.cptr, C.c_size_t(dB.nbytes))
            lu.solve(X)
        else:
            lu.factor(A)
            LA.solve_refine(A, dB, lu, tol=1e-12, out=X)
        _lib.call("pvl_synchronize")
        split = {}
        for k in ("lu_panel", "lu_gemm", "lu_trsm", "lu_laswp", "lu_trsv", "lu_diag_inverse"):
            ms, cnt, work = prof_get(lib, f"{k}_{sfx}")
            split[k] = {"ms": round(ms, 3), "launches": cnt}
            if k == "lu_gemm":
                gemm_tf = work / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        for k in ("residual_f64",):
            ms, cnt, _ = prof_get(lib, k)
            if cnt:
                split[k] = {"ms": round(ms, 3), "launches": cnt}
        lib.pvl_prof_enable(0)
        achieved = alg_flops / tw / 1e12
        roof[name] = {
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4),
            "note": f"whole solve: 2/3 n^3 + 2 n^2 nrhs = {alg_flops:.4g} flop (one factorisation, n = {n}, "
                    f"{nrhs} right-hand sides) over its wall time; refinement's extra residuals not counted",
            "trailing_update_gemm": {"achieved": round(gemm_tf, 3), "frac": round(gemm_tf / peak, 4)},
            "kernel_ms": split,
        }
        log(f"C4 {name}: {split}; gemm {gemm_tf:.1f} TFLOP/s")
    bmax = np.max(np.abs(B), axis=0)
    res64 = float(np.max(np.max(np.abs(LA.residual(A, X64, dB).numpy()), axis=0) / bmax))
    rel = lambda a, b: float(np.max(np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)))  # noqa: E731
    out = {
        "n": n,
        "assembly_fp64_ms": round(asm_ms, 3),
        "assembly_alg_tflops": round(asm_pairs * FLOP_PER_PAIR / (asm_ms * 1e-3) / 1e12, 3),
        "fp64_lu_solve_s": round(t64, 4),
        "fp32_mfma_lu_ir_s": round(t32, 4),
        "ir_corrections": iters,
        "ir_target_rel_residual": 1e-12,
        "ir_final_rel_residual": float(np.max(relres)),
        "fp64_lu_rel_residual": res64,
        "rel_err_fp32ir_vs_fp64gpu": rel(x32, x64),
        "fp32_mfma_lu_nopiv_ir_s": round(t32n, 4),
        "ir_nopiv_corrections": iters_n,
        "ir_nopiv_final_rel_residual": float(np.max(relres_n)),
        "rel_err_fp32nopiv_ir_vs_fp64gpu": rel(x32n, x64),
        "deflation_c": c,
        "speedup_fp32ir_over_fp64": round(t64 / t32, 3),
        "speedup_fp32nopiv_ir_over_fp64": round(t64 / t32n, 3),
        "roofline": roof,
    }
    if do_cpu:
        import scipy.linalg as la

        Ad = A.numpy()
        t0 = time.perf_counter()
        xr_raw = la.lu_solve(la.lu_factor(Ad.copy(), overwrite_a=True, check_finite=False), B, check_finite=False)
        tcpu = time.perf_counter() - t0
        xr = mean_free(xr_raw)
        out.update({
            "cpu_lapack_fp64_s": round(tcpu, 3),
            "cpu_cores": os.cpu_count(),
            "rel_err_fp64gpu_vs_lapack": rel(x64, xr),
            "rel_err_fp32ir_vs_lapack": rel(x32, xr),
            "rel_err_fp32nopiv_ir_vs_lapack": rel(x32n, xr),
            "lapack_rel_residual": float(np.max(np.max(np.abs(B - Ad @ xr_raw), axis=0) / bmax)),
        })
        log(f"C4 scipy LAPACK FP64 on {os.cpu_count()} cores: {tcpu:.2f} s")
    return out


def run_c5(train_steps: int = 2000, eval_steps: int = 300, seed: int = 7) -> dict:
    """Learned warm start: pitch + heave tapered wing, 2,048 panels, 300-row free wake.

    The network is trained on ``train_steps`` FP64-LU steps from kinematics seeds 100-107
    (disjoint from the evaluation seed), then the evaluation run is repeated with the plain
    FP64 LU and with FP32 GMRES in FP64 refinement started from zero, from the previous
    step's circulation, from its linear extrapolation, from the extrapolation corrected by the
    network, and from the minimal-residual combination of the last three circulations without
    and with the network's prediction as a further direction.  Errors: per step, max|dG| /
    max|G| against the FP64 LU run."""
    from pyvl_amd import learned as L

    t0 = time.perf_counter()
    X, y = L.generate_training_data(train_steps, seeds=tuple(range(100, 108)))
    t_gen = time.perf_counter() - t0
    norm = L.Normaliser(X, y)
    net = L.MLP(X.shape[1], 256, 4, seed=42)
    t0 = time.perf_counter()
    losses = L.train(net, X, y, norm, epochs=4, batch=8192, lr=2e-3, log=log)
    t_train = time.perf_counter() - t0
    log(f"C5 data: {X.shape[0]} samples in {t_gen:.1f} s; training {t_train:.1f} s, losses {losses}")
    runs = {"lu64": L.run_c5(seed, eval_steps, solve="lu64")}
    for w in ("none", "prev", "extrap", "mlp", "proj", "proj_mlp"):
        runs["gmres_" + w] = L.run_c5(seed, eval_steps, solve="gmres", warm=w, mlp=net, norm=norm)
    ref = runs["lu64"]["gamma"]
    out = {"panels": int(ref.shape[1]), "train_samples": int(X.shape[0]), "train_data_s": round(t_gen, 2),
           "train_s": round(t_train, 2), "train_loss_first_last": [losses[0], losses[-1]]}
    for k, r in runs.items():
        err = np.max(np.abs(r["gamma"] - ref), axis=1) / np.max(np.abs(ref), axis=1)
        st = r["gmres"]
        out[k] = {
            "run_s": round(float(np.sum(r["step_s"])), 4),
            "solve_ms_mean": round(1e3 * float(np.mean(r["solve_s"])), 4),
            "gmres_steps_mean": round(float(np.mean([s[1] for s in st])), 3) if st else None,
            "corrections_mean": round(float(np.mean([s[0] for s in st])), 3) if st else None,
            "start_rel_residual_median": float(np.median([s[2] for s in st[2:]])) if st else None,
            "fallbacks": r["fallbacks"],
            "max_rel_err_vs_fp64_lu": float(np.max(err)),
        }
        log(f"C5 {k}: {out[k]}")
    r0 = {k: out[k]["start_rel_residual_median"] for k in runs if k != "lu64"}
    out["start_residual_gain_network_over_extrapolation"] = round(r0["gmres_extrap"] / r0["gmres_mlp"], 3)
    out["start_residual_gain_network_over_projection"] = round(r0["gmres_proj"] / r0["gmres_proj_mlp"], 3)
    out["run_speedup_gmres_proj_over_fp64_lu"] = round(out["lu64"]["run_s"] / out["gmres_proj"]["run_s"], 3)
    out["run_speedup_gmres_mlp_over_fp64_lu"] = round(out["lu64"]["run_s"] / out["gmres_mlp"]["run_s"], 3)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the host baselines")
    ap.add_argument("--headline-only", action="store_true", help="only the FP64 free-wake workload")
    ap.add_argument("--solve-only", action="store_true",
                    help="only the 16k-panel closed-body solve (FP64 LU, FP32 LU + FP64 refinement); prints its "
                         "own JSON line, not the headline one")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device-to-device (a plain copy kernel of known size, the reference "
                         "for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}); "
              "the workload would shard by evaluation point, which is a design note, not code", file=sys.stderr)
        return 2
    from pyvl_amd import _lib

    _lib.call("pvl_init", 0)
    if args.solve_only:
        print(json.dumps({"workload": "16k-panel closed-body solve, 16 RHS", **run_c4(not args.no_cpu)}), flush=True)
        return 0
    if args.copy_calibration:
        from pyvl_amd import DeviceArray

        src, dst = DeviceArray((1 << 27,)), DeviceArray((1 << 27,))  # 1 GiB each
        _lib.call("pvl_memset_zero", src.cptr, C.c_size_t(src.nbytes))
        _lib.call("pvl_memcpy_d2d", dst.cptr, src.cptr, C.c_size_t(src.nbytes))
        _lib.call("pvl_synchronize")
        log("copy calibration: 1 GiB device-to-device copy done")
        src.free()
        dst.free()
    head = run_c3("free", 0, args.steps, args.warmup, profile=True)
    extra = {"c3_grow_400_steps_s": round(head["grow_400_s"], 3)}
    extra["c3_kernel_breakdown_ms_per_step"] = head["breakdown"]
    cpu = None
    if not args.headline_only:
        f32 = run_c3("free", 1, args.steps, args.warmup, profile=False)
        extra["c3_free_wake_fp32_induction_steps_per_s"] = round(f32["steps_per_s"], 3)
        extra["c3_free_wake_fp32_induction_grow_400_s"] = round(f32["grow_400_s"], 3)
        ref = run_c3("reference", 0, args.steps, args.warmup, profile=False)
        extra["c3_reference_semantics_steps_per_s"] = round(ref["steps_per_s"], 3)
        extra["c3_reference_semantics_grow_400_s"] = round(ref["grow_400_s"], 3)
        extra["solve_16k"] = run_c4(not args.no_cpu)
        extra["learned_warm_start"] = run_c5()
        if not args.no_cpu:
            extra["cpu_reference_semantics"] = cpu_reference_step(ref, 2)
            log(f"CPU reference semantics: {extra['cpu_reference_semantics']}")
    if not args.no_cpu:
        cpu = cpu_free_wake_step(head)
        log(f"CPU free wake: {cpu}")
    if not args.headline_only:
        extra["c3_free_wake_fp32_induction_vs_fp64"] = fp32_vs_fp64_step(f32["sol"], head["sol"])
        log(f"C3 FP32 induction against FP64 from the same state: {extra['c3_free_wake_fp32_induction_vs_fp64']}")
    line = {
        "metric": "wake steps/s (1k bound+25k wake panels)",
        "value": round(head["steps_per_s"], 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "unsteady free wake: flat wing 64x16 = 1,024 bound panels (span 8, chord 1), heave 0.1c, "
                        "k=0.3, dt=c/16U, 400-row wake = 25,600 wake panels (26,065 nodes, 51,664 segments), "
                        "tol 1e-6, FP64; timed steps run with the wake full",
            "bound_panels": 1024,
            "wake_panels": 25600,
            "line_rows": 400,
        },
        "roofline": roofline_from(head["prof"]),
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
