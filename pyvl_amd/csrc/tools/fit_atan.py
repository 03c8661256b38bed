"""Fits the minimax polynomial of the library's FP64/FP32 arctangent.

atan(a) ~= a + a * s * P(s), s = a^2, 0 <= a <= 1, minimising the maximum relative error
of atan (Lawson's iteratively reweighted least squares on Chebyshev nodes, in 60-digit
arithmetic), then checks the double/float Horner evaluation against mpmath.  Prints the
coefficients as C literals.  Usage: python fit_atan.py DEGREE_P [float]
"""
import sys

import mpmath as mp
import numpy as np

mp.mp.dps = 60


def fit(deg, n_nodes=400, iters=60):
    nodes = [(1 - mp.cos(mp.pi * (k + mp.mpf(0.5)) / n_nodes)) / 2 for k in range(n_nodes)]  # s in (0,1)
    g = []
    w_rel = []
    for s in nodes:
        a = mp.sqrt(s)
        g.append((mp.atan(a) - a) / (a * s))
        w_rel.append(a * s / mp.atan(a))  # converts P error into relative atan error
    lw = [mp.mpf(1)] * n_nodes
    coef = None
    for _ in range(iters):
        # weighted least squares: minimise sum lw * (w_rel*(P - g))^2
        A = mp.matrix(n_nodes, deg + 1)
        b = mp.matrix(n_nodes, 1)
        for i, s in enumerate(nodes):
            sw = mp.sqrt(lw[i]) * w_rel[i]
            p = mp.mpf(1)
            for j in range(deg + 1):
                A[i, j] = sw * p
                p *= s
            b[i] = sw * g[i]
        coef = mp.lu_solve(A.T * A, A.T * b)
        err = []
        for i, s in enumerate(nodes):
            p = mp.mpf(0)
            for j in reversed(range(deg + 1)):
                p = p * s + coef[j]
            err.append(abs(w_rel[i] * (p - g[i])))
        tot = sum(lw[i] * err[i] for i in range(n_nodes))
        lw = [lw[i] * err[i] / tot for i in range(n_nodes)]
    return [coef[j] for j in range(deg + 1)], max(err)


def check(coef, dtype):
    rng = np.random.default_rng(0)
    a = np.concatenate([np.linspace(0, 1, 200001), rng.uniform(0, 1, 200000), 10.0 ** -rng.uniform(0, 12, 20000)])
    a = a.astype(dtype)
    c = [dtype(float(x)) for x in coef]
    s = a * a
    p = np.full_like(a, c[-1])
    for x in reversed(c[:-1]):
        p = p * s + x  # numpy has no fma; the device uses fma (slightly better)
    r = a + a * (s * p)
    worst = 0.0
    for ai, ri in zip(a[::50], r[::50]):
        t = mp.atan(mp.mpf(float(ai)))
        if t != 0:
            worst = max(worst, float(abs((mp.mpf(float(ri)) - t) / t)))
    return worst


if __name__ == "__main__":
    deg = int(sys.argv[1])
    dtype = np.float32 if len(sys.argv) > 2 and sys.argv[2] == "float" else np.float64
    coef, err = fit(deg)
    print(f"// degree {deg}: minimax relative error of the polynomial {float(err):.3e}, "
          f"{dtype.__name__} Horner {check(coef, dtype):.3e}")
    for x in coef:
        print(f"    {mp.nstr(x, 20, min_fixed=-1, max_fixed=-1)},")
