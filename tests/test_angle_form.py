"""CPU pins of the angle arithmetic used by the induction kernels (induction_core.hpp):
the single-angle identity that replaces the reference's difference of two atan2 values, and
the minimax arctangent polynomials (coefficients parsed from the kernel source)."""

import re
from pathlib import Path

import numpy as np

SRC = (Path(__file__).resolve().parents[1] / "pyvl_amd" / "csrc" / "induction_core.hpp").read_text()


def _coeffs(fn):
    body = SRC[SRC.index(fn):]
    body = body[: body.index("return")]
    vals = [float(x.rstrip("f")) for x in re.findall(r"(-?\d\.\d+e-?\d+f?)", body)]
    return vals  # highest degree first, as in the Horner code


def _atan_poly(a, c):
    s = a * a
    p = np.full_like(a, c[0])
    for x in c[1:]:
        p = p * s + x
    return a + a * s * p


def test_single_angle_identity():
    rng = np.random.default_rng(0)
    t1 = rng.normal(0, 3, 200000)
    t2 = t1 - np.abs(rng.normal(0, 2, t1.size)) - 1e-3  # t2 - t1 = -len < 0 as for a segment
    t2[::2] = t1[::2] + np.abs(rng.normal(0, 2, t1.size // 2)) + 1e-3  # and the other sign
    h = np.abs(rng.normal(0, 1, t1.size)) + 1e-6
    two = np.arctan2(t2, h) - np.arctan2(t1, h)
    one = np.arctan2((t2 - t1) * h, h * h + t1 * t2)
    np.testing.assert_allclose(one, two, rtol=1e-12, atol=1e-13)
    # far field: the single form keeps relative accuracy where the difference cancels
    import mpmath as mp

    mp.mp.dps = 50
    worst_one = worst_two = 0.0
    for t1f, t2f, hf in [(1000.0, 999.999, 1000.0), (35.0, 34.93, 12.5), (-400.0, -400.06, 3.0)]:
        exact = mp.atan2(t2f, hf) - mp.atan2(t1f, hf)
        one = np.arctan2((np.float64(t2f) - t1f) * hf, hf * hf + t1f * t2f)
        two = np.arctan2(t2f, hf) - np.arctan2(t1f, hf)
        worst_one = max(worst_one, float(abs((one - exact) / exact)))
        worst_two = max(worst_two, float(abs((two - exact) / exact)))
    assert worst_one < 1e-15 < worst_two


def test_segment_length_form():
    """For a segment, t2 - t1 = e . (r1 - r2) = -len exactly; the kernels use -len h as the
    sine term.  Against 50-digit arithmetic on random segments and points (near and far): the
    -len form's median and worst errors are below those of the rounded difference t2 - t1
    (pair by pair either may win by a few ulp); what error remains comes from h^2 (the
    reference's |dr|^2 - t^2 cancellation, kept on purpose), which both forms share."""
    import mpmath as mp

    mp.mp.dps = 50
    rng = np.random.default_rng(3)
    errs = []
    for k in range(400):
        r1 = rng.normal(size=3)
        d = rng.normal(size=3) * 10.0 ** rng.uniform(-2, 0)
        r2 = r1 + d
        p = r1 + rng.normal(size=3) * 10.0 ** rng.uniform(-1, 2)
        length = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        e = d / length
        d1, d2 = p - r1, p - r2
        t1 = e[0] * d1[0] + e[1] * d1[1] + e[2] * d1[2]
        t2 = e[0] * d2[0] + e[1] * d2[1] + e[2] * d2[2]
        h2 = ((d1 @ d1 - t1 * t1) + (d2 @ d2 - t2 * t2)) / 2.0
        h = np.sqrt(h2)
        got = np.arctan2(-length * h, h2 + t1 * t2)
        old = np.arctan2((t2 - t1) * h, h2 + t1 * t2)
        P = [mp.mpf(float(x)) for x in p]
        A = [mp.mpf(float(x)) for x in r1]
        B = [mp.mpf(float(x)) for x in r2]
        L = mp.sqrt(sum((B[i] - A[i]) ** 2 for i in range(3)))
        E = [(B[i] - A[i]) / L for i in range(3)]
        T1 = sum(E[i] * (P[i] - A[i]) for i in range(3))
        T2 = sum(E[i] * (P[i] - B[i]) for i in range(3))
        H = mp.sqrt(sum((P[i] - A[i]) ** 2 for i in range(3)) - T1 * T1)
        exact = mp.atan2(T2, H) - mp.atan2(T1, H)
        errs.append((float(abs((got - exact) / exact)), float(abs((old - exact) / exact))))
    errs = np.array(errs)
    assert np.median(errs[:, 0]) < np.median(errs[:, 1]) and errs[:, 0].max() < errs[:, 1].max()
    assert np.median(errs[:, 0]) < 1e-15


def test_fp64_atan_polynomial():
    c = _coeffs("atan_poly64")
    assert len(c) == 17
    a = np.concatenate([np.linspace(0, 1, 400001), 10.0 ** -np.linspace(0, 15, 2000)])
    got = _atan_poly(a, c)
    ref = np.arctan(a)
    nz = ref > 0
    assert np.max(np.abs(got[nz] - ref[nz]) / ref[nz]) < 1.3e-15


def test_fp32_atan_polynomial():
    c = _coeffs("atan_poly32")
    assert len(c) == 8
    a = np.linspace(0, 1, 200001).astype(np.float32)
    got = _atan_poly(a, [np.float32(x) for x in c]).astype(np.float64)
    ref = np.arctan(a.astype(np.float64))
    nz = ref > 0
    assert np.max(np.abs(got[nz] - ref[nz]) / ref[nz]) < 3e-7


def test_octant_reduction():
    """atan2 from atan on [0, 1]: pi/2 - . when |y| > |x|, pi - . when x < 0, sign of y."""
    rng = np.random.default_rng(1)
    y = rng.normal(size=100000)
    x = rng.normal(size=100000)
    ax, ay = np.abs(x), np.abs(y)
    t = np.arctan(np.minimum(ax, ay) / np.maximum(ax, ay))
    t = np.where(ay > ax, np.pi / 2 - t, t)
    t = np.where(x < 0, np.pi - t, t)
    t = np.copysign(t, y)
    np.testing.assert_allclose(t, np.arctan2(y, x), rtol=0, atol=1e-15)
