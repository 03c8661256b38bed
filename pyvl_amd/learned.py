"""Learned warm start for the per-step circulation solve.

A 4-layer MLP (width 256, BF16 products with FP32 accumulation on MFMA, FP32 master weights,
Adam) is trained on the GPU, by libpvl_hip.so's own forward/backward kernels, on pairs
that the solver itself generates: per panel, the step's features and how far its
circulation departs from the linear extrapolation of the two previous steps.  At solve time
the extrapolation plus the predicted departure is the starting point of FP32 GMRES inside FP64 iterative refinement
(``pvl_gmres_refine``); a solve that does not converge falls back to the FP64 LU.  There is
no reference counterpart: the reference factors every step with LAPACK
(python/pyvl/solver.py:223-227).

Per-panel features (``FEATURES``): the step's right-hand side before and after the wake's
explicit correction (free stream + body motion, and the wake-induced normal velocity), its
first and second differences in time, the previous circulation and its last change, the
panel's position on the wing and its normal, whether it borders the trailing edge, and the
body's pitch angle, rate and acceleration and heave rate and acceleration.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .device import DeviceArray

FEATURES = (
    "rhs_body", "rhs_wake", "rhs_total", "d_rhs_total", "dd_rhs_total", "gamma_prev", "d_gamma_prev", "x_rel",
    "y_rel", "chord_frac", "n_x", "n_z", "trailing_edge", "pitch", "pitch_rate", "pitch_acc", "heave_rate",
    "heave_acc",
)


class MLP:
    """Device MLP: d_in -> width x (n_layers - 1) -> 1 with SiLU activations."""

    def __init__(self, d_in: int, width: int = 256, n_layers: int = 4, seed: int = 42):
        h = C.c_void_p()
        _lib.call("pvl_mlp_create", C.byref(h), C.c_int(d_in), C.c_int(width), C.c_int(n_layers))
        self._h = h
        self.d_in, self.width, self.n_layers = d_in, width, n_layers
        self.dims = [d_in] + [width] * (n_layers - 1) + [1]
        n = C.c_long()
        _lib.call("pvl_mlp_num_params", h, C.byref(n))
        self.n_params = n.value
        self.set_params(self.init_params(seed))

    def init_params(self, seed: int) -> np.ndarray:
        """He-uniform weights (SiLU), zero biases; layout: per layer W (out x in) then b."""
        rng = np.random.default_rng(seed)
        parts = []
        for l in range(self.n_layers):
            fan_in, fan_out = self.dims[l], self.dims[l + 1]
            lim = np.sqrt(6.0 / fan_in) if l < self.n_layers - 1 else np.sqrt(3.0 / fan_in)
            parts.append(rng.uniform(-lim, lim, (fan_out, fan_in)).ravel())
            parts.append(np.zeros(fan_out))
        p = np.concatenate(parts).astype(np.float32)
        assert p.size == self.n_params
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        _lib.call("pvl_mlp_set_params", self._h, p.ctypes.data_as(C.c_void_p))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        _lib.call("pvl_mlp_get_params", self._h, p.ctypes.data_as(C.c_void_p))
        return p

    def unpack(self, p: np.ndarray):
        """[(W, b)] per layer from a flat parameter vector."""
        out, off = [], 0
        for l in range(self.n_layers):
            fi, fo = self.dims[l], self.dims[l + 1]
            w = p[off:off + fi * fo].reshape(fo, fi)
            off += fi * fo
            b = p[off:off + fo]
            off += fo
            out.append((w, b))
        return out

    def forward(self, x: DeviceArray, out: DeviceArray | None = None) -> DeviceArray:
        y = out if out is not None else DeviceArray((x.shape[0],), np.float32)
        _lib.call("pvl_mlp_forward", self._h, x.cptr, y.cptr, C.c_int(x.shape[0]))
        return y

    def gradient(self, x: DeviceArray, t: DeviceArray):
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        _lib.call("pvl_mlp_gradient", self._h, x.cptr, t.cptr, C.c_int(x.shape[0]), g.ctypes.data_as(C.c_void_p),
                  C.byref(loss))
        return g, float(loss.value)

    def train_step(self, x: DeviceArray, t: DeviceArray, lr: float, wd: float = 0.0, want_loss: bool = False):
        loss = C.c_float()
        _lib.call("pvl_mlp_train_step", self._h, x.cptr, t.cptr, C.c_int(x.shape[0]), C.c_float(lr),
                  C.c_float(0.9), C.c_float(0.999), C.c_float(1e-8), C.c_float(wd),
                  C.byref(loss) if want_loss else None)
        return float(loss.value) if want_loss else None

    def __del__(self):
        try:
            if getattr(self, "_h", None) is not None and _lib.alive():
                _lib._LIB.pvl_mlp_destroy(self._h)
                self._h = None
        except Exception:
            pass


def gmres_refine(a64: DeviceArray, a32: DeviceArray, b: DeviceArray, x: DeviceArray, tol: float = 1e-12,
                 m: int = 32, inner_tol: float = 1e-7, max_outer: int = 6):
    """x <- solution of A x = b (x holds the start) by FP32 GMRES in FP64 refinement, to
    max|b - A x| <= tol max|b|.  Returns (converged, corrections, gmres_steps, start
    relative residual, final relative residual)."""
    stats = (C.c_int * 3)()
    res = (C.c_double * 2)()
    st = _lib.load().pvl_gmres_refine(C.c_int(a64.shape[0]), a64.cptr, a32.cptr, b.cptr, x.cptr, C.c_double(tol),
                                      C.c_int(m), C.c_double(inner_tol), C.c_int(max_outer), stats, res)
    if st not in (_lib.PVL_OK, _lib.PVL_ERR_NOTCONV):
        _lib.check(st, "pvl_gmres_refine")
    return st == _lib.PVL_OK, stats[0], stats[1], res[0], res[1]


class FeatureBuilder:
    """Per-panel features of the current step (host side; N x len(FEATURES) float32) and the
    extrapolated circulation the network corrects.

    ``base()``: 2 G(t-1) - G(t-2) once two previous circulations exist, G(t-1) after one
    step, None before.  The network predicts G(t) - base(); the second difference of the
    right-hand side (``dd_rhs_total``) and the body's accelerations drive that remainder.
    """

    def __init__(self, solver, pos0: np.ndarray, te_surfaces: np.ndarray, chord_root: float, span: float):
        self.solver = solver
        cps0 = solver.mesh.surface_average_vec3(pos0)
        self.x_rel = cps0[:, 0] / chord_root
        self.y_rel = cps0[:, 1] / (0.5 * span)
        le = np.zeros_like(self.x_rel)
        # local chord fraction: position between the panel row's leading and trailing x
        ys = np.round(cps0[:, 1], 9)
        for y in np.unique(ys):
            sel = ys == y
            xs = cps0[sel, 0]
            lo, hi = xs.min(), xs.max()
            le[sel] = (xs - lo) / max(hi - lo, 1e-12)
        self.chord_frac = le
        self.te = np.zeros(cps0.shape[0])
        self.te[te_surfaces] = 1.0
        self.totals: list[np.ndarray] = []   # the last two steps' rhs, newest last
        self.gammas: list[np.ndarray] = []   # the last three steps' circulations, newest last

    @property
    def history(self) -> list:
        return self.gammas[::-1]

    @property
    def prev_gamma(self):
        return self.gammas[-1] if self.gammas else None

    def base(self):
        if len(self.gammas) >= 2:
            return 2.0 * self.gammas[-1] - self.gammas[-2]
        return self.prev_gamma

    def build(self, rhs_body: np.ndarray, rhs_total: np.ndarray, normals: np.ndarray, kin: tuple) -> np.ndarray:
        n = rhs_body.size
        zero = np.zeros(n)
        t1 = self.totals[-1] if self.totals else rhs_total
        t2 = self.totals[-2] if len(self.totals) >= 2 else t1
        g1 = self.gammas[-1] if self.gammas else zero
        g2 = self.gammas[-2] if len(self.gammas) >= 2 else g1
        pitch, pitch_rate, pitch_acc, heave_rate, heave_acc = kin
        f = np.stack([
            rhs_body, rhs_total - rhs_body, rhs_total, rhs_total - t1, rhs_total - 2.0 * t1 + t2, g1, g1 - g2,
            self.x_rel, self.y_rel, self.chord_frac, normals[:, 0], normals[:, 2], self.te,
            np.full(n, np.sin(pitch)), np.full(n, pitch_rate), np.full(n, pitch_acc), np.full(n, heave_rate),
            np.full(n, heave_acc),
        ], axis=1)
        return f.astype(np.float32)

    def commit(self, rhs_total, gamma: np.ndarray) -> None:
        if rhs_total is not None:
            self.totals = (self.totals + [rhs_total.copy()])[-2:]
        self.gammas = (self.gammas + [gamma.copy()])[-3:]


class Projector:
    """Minimal-residual start over a few candidate vectors: x0 = V c with
    c = argmin ||b - A V c||_2 (V = the last solutions and, optionally, the network's
    prediction).  A V is one FP64 pass over A on the device (``pvl_residual`` with B = 0);
    the k x k least-squares problem is solved on the host.  The prediction thereby enters
    as a search direction with its best coefficient, so a poor prediction cannot make the
    start worse than the previous solutions alone (in the 2-norm of the residual)."""

    def __init__(self, n: int, k_max: int = 4):
        self.n, self.k_max = n, k_max
        self.V = DeviceArray((n, k_max))
        self.Z = DeviceArray.zeros((n, k_max))
        self.W = DeviceArray((n, k_max))

    def start(self, a: DeviceArray, b_host: np.ndarray, vectors) -> np.ndarray | None:
        vecs = [v for v in vectors if v is not None]
        if not vecs:
            return None
        k = len(vecs)
        if k > self.k_max:
            raise ValueError(f"at most {self.k_max} vectors")
        V = np.ascontiguousarray(np.stack(vecs, axis=1), np.float64)
        Vd, Zd, Wd = (d.view(0, (self.n, k)) for d in (self.V, self.Z, self.W))
        Vd.copy_from(V)
        _lib.call("pvl_residual", C.c_int(self.n), a.cptr, Vd.cptr, Zd.cptr, Wd.cptr, C.c_int(k))
        AV = -Wd.numpy()
        c, *_ = np.linalg.lstsq(AV, b_host, rcond=1e-12)
        return V @ c


class ProjectionStart:
    """Default start of ``UnsteadySolver(solve="gmres")``: the minimal-residual combination
    (:class:`Projector`) of the last ``k`` circulations the solver produced (Fischer's
    projection for successive right-hand sides); zero at the first step.  Called by the
    solver before each solve, when its ``d_x`` still holds the previous step's circulation.
    A jump in the iteration count (a restored state) restarts the history from the
    circulation the solver then holds."""

    def __init__(self, n: int, k: int = 3):
        self.proj = Projector(n, k_max=k)
        self.k = k
        self.hist: list[np.ndarray] = []
        self.last: int | None = None

    def __call__(self, sol) -> np.ndarray | None:
        it = int(sol.iteration)
        if it == 0:
            self.hist, self.last = [], 0
            return None
        if self.last is None or it != self.last + 1:
            self.hist = []
        self.last = it
        self.hist = (self.hist + [sol.d_x.numpy()])[-self.k:]
        return self.proj.start(sol.d_A, sol.d_b.numpy(), self.hist)


class Normaliser:
    """Feature and target standardisation from training statistics."""

    def __init__(self, X: np.ndarray, y: np.ndarray):
        self.mu = X.mean(axis=0)
        self.sd = X.std(axis=0)
        self.sd[self.sd < 1e-12] = 1.0
        self.ymu = float(y.mean())
        self.ysd = float(y.std()) or 1.0

    def x(self, X: np.ndarray) -> np.ndarray:
        return ((X - self.mu) / self.sd).astype(np.float32)

    def y(self, y: np.ndarray) -> np.ndarray:
        return ((y - self.ymu) / self.ysd).astype(np.float32)

    def y_inv(self, z: np.ndarray) -> np.ndarray:
        return z.astype(np.float64) * self.ysd + self.ymu


# ---------------------------------------------------------------------------------------
# The pitch + heave tapered-wing workload (2,048 panels, free wake): data, training, use.
def _kinematics(frame, t):
    """Pitch angle, rate and acceleration; heave rate and acceleration."""
    th, thd, h, hd = frame.motion(t)
    w2 = frame.omega * frame.omega
    return float(th), float(thd), float(-w2 * th), float(hd), float(-w2 * h)


def run_c5(kin_seed: int, steps: int, solve: str = "lu64", warm: str = "none", mlp: MLP | None = None,
           norm: Normaliser | None = None, record: bool = False, rows: int | None = None, pre_solve=None):
    """Run the workload; returns per-step circulations, times and GMRES statistics.

    solve "lu64": FP64 LU every step; "gmres": FP32 GMRES in FP64 refinement (LU fallback)
    started from zero ("none"), the previous step's circulation ("prev"), its linear
    extrapolation 2 G(t-1) - G(t-2) ("extrap"), that extrapolation plus the network's
    predicted remainder ("mlp"), the minimal-residual combination of the last three
    circulations ("proj", :class:`Projector`) or of those and the network's prediction
    ("proj_mlp"); steps without enough history start as "prev" does.
    ``pre_solve(solver, it)`` is called after the step's system is assembled.  ``solve_s``:
    features + prediction + solve; ``step_s``: the whole step.
    """
    import time

    from .solver import C5, make_c5_solver

    if warm in ("mlp", "proj_mlp") and (mlp is None or norm is None):
        raise ValueError("warm='mlp' needs a trained network and its normaliser")
    sol = make_c5_solver(kin_seed, solve=solve, rows=rows)
    n = sol.mesh.n_surfaces
    if sol.d_b0 is None:
        sol.d_b0 = DeviceArray((n,))
    adj = np.asarray(sol.wake.adjacent_surfaces).reshape(-1, 2)
    te = np.unique(adj[(adj != 0x7FFFFFFF)].astype(np.int64))
    fb = FeatureBuilder(sol, sol.p0, te, C5["chord"], C5["span"])
    dX = DeviceArray((n, len(FEATURES)), np.float32)
    dY = DeviceArray((n,), np.float32)
    gammas, times, steps_s, feats = [], [], [], []
    x0 = {"v": None}
    sol.warm_start = lambda s: x0["v"]
    need_feats = record or warm in ("mlp", "proj_mlp")
    proj = Projector(n) if warm in ("proj", "proj_mlp") else None
    _lib.call("pvl_synchronize")
    for it in range(steps):
        ts = time.perf_counter()
        t = it * sol.dt
        sol.assemble(t)
        if pre_solve is not None:
            pre_solve(sol, it)
        _lib.call("pvl_synchronize")
        t0 = time.perf_counter()
        if need_feats:
            rb, rt, nr = sol.d_b0.numpy(), sol.d_b.numpy(), sol.d_nrm.numpy()
            f = fb.build(rb, rt, nr, _kinematics(sol.frame, t))
        pred = None
        if warm in ("mlp", "proj_mlp") and len(fb.gammas) >= 2:
            dX.copy_from(norm.x(f))
            mlp.forward(dX, dY)
            pred = fb.base() + norm.y_inv(dY.numpy())
        if warm == "mlp":
            x0["v"] = pred if pred is not None else fb.base()
        elif warm == "extrap":
            x0["v"] = fb.base()
        elif warm == "prev":
            x0["v"] = fb.prev_gamma
        elif warm in ("proj", "proj_mlp"):
            x0["v"] = proj.start(sol.d_A, sol.d_b.numpy(), list(fb.history) + [pred])
        else:
            x0["v"] = None
        sol.solve_system()
        _lib.call("pvl_synchronize")
        times.append(time.perf_counter() - t0)
        g = sol.d_x.numpy()
        gammas.append(g)
        if record:
            feats.append(f)
        fb.commit(rt if need_feats else None, g)
        sol.wake.update(t, sol.mesh, sol.d_pos, sol.d_x, sol.flow)
        sol.iteration += 1
        _lib.call("pvl_synchronize")
        steps_s.append(time.perf_counter() - ts)
    return {
        "gamma": np.array(gammas), "solve_s": np.array(times), "step_s": np.array(steps_s),
        "features": np.array(feats) if record else None, "gmres": list(sol.gmres_stats), "fallbacks": sol.fallbacks,
    }


def generate_training_data(n_steps: int = 2000, seeds=tuple(range(100, 108)), rows: int | None = None,
                           solve: str = "gmres"):
    """Solver-generated training pairs from runs of the workload with kinematics seeds
    disjoint from the evaluation seed: per panel and step, the features and the
    circulation's departure from the linear extrapolation 2 G(t-1) - G(t-2) (steps 0 and 1,
    without two previous circulations, are left out).  ``solve`` "gmres": the refined GMRES
    started by the minimal-residual projection (the same circulations as the FP64 LU to
    ~1e-12, at a fraction of the cost); "lu64": the FP64 LU."""
    per = int(np.ceil(n_steps / len(seeds)))
    X, y = [], []
    left = n_steps
    for s in seeds:
        k = min(per, left)
        if k <= 0:
            break
        r = run_c5(s, k, solve=solve, warm="proj" if solve == "gmres" else "none", record=True, rows=rows)
        g = r["gamma"]
        X.append(r["features"][2:].reshape(-1, len(FEATURES)))
        y.append((g[2:] - 2.0 * g[1:-1] + g[:-2]).reshape(-1))
        left -= k
    return np.concatenate(X), np.concatenate(y)


def train(mlp: MLP, X: np.ndarray, y: np.ndarray, norm: Normaliser, epochs: int = 4, batch: int = 8192,
          lr: float = 2e-3, seed: int = 42, log=None):
    """Adam on shuffled mini-batches, cosine learning-rate decay; returns per-epoch losses
    (mean squared error of the standardised target)."""
    rng = np.random.default_rng(seed)
    Xn, yn = norm.x(X), norm.y(y)
    n = Xn.shape[0] // batch * batch
    dX = DeviceArray((n, Xn.shape[1]), np.float32)
    dy = DeviceArray((n,), np.float32)
    total = epochs * (n // batch)
    k = 0
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(Xn.shape[0])[:n]
        dX.copy_from(Xn[perm])
        dy.copy_from(yn[perm])
        ep_loss = []
        for b in range(n // batch):
            cur = lr * 0.5 * (1.0 + np.cos(np.pi * k / total))
            want = b % 50 == 0 or b == n // batch - 1
            l = mlp.train_step(dX.view(b * batch * Xn.shape[1], (batch, Xn.shape[1])), dy.view(b * batch, (batch,)),
                               cur, wd=1e-5, want_loss=want)
            if want:
                ep_loss.append(l)
            k += 1
        losses.append(float(np.mean(ep_loss)))
        if log:
            log(f"epoch {ep}: loss {losses[-1]:.3e}")
    return losses
