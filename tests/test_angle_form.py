"""CPU pins of the arithmetic used by the induction kernels (induction_core.hpp): the
single-angle identity that replaces the reference's difference of two atan2 values, the
distance from the cross product, the refined reciprocal square root, and the minimax
arctangent polynomials (coefficients parsed from the kernel source)."""

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
    (pair by pair either may win by a few ulp); both use the reference's h^2 here, so what
    error remains comes from its |dr|^2 - t^2 cancellation (test_cross_product_distance_form
    covers the cross-product h^2 of the wake's self-induction)."""
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


def _exact_k(d1, e, length):
    """K from the rounded inputs both forms share (dr1 = p - r1, the unit direction e and the
    length, as float64), every operation after them in 50-digit arithmetic (mpmath)."""
    import mpmath as mp

    D = [mp.mpf(float(x)) for x in d1]
    E = [mp.mpf(float(x)) for x in e]
    L = mp.mpf(float(length))
    T1 = sum(E[i] * D[i] for i in range(3))
    T2 = T1 - L
    C = [D[1] * E[2] - D[2] * E[1], D[2] * E[0] - D[0] * E[2], D[0] * E[1] - D[1] * E[0]]
    H2 = sum(c * c for c in C)
    H = mp.sqrt(H2)
    M = (mp.atan2(T2, H) - mp.atan2(T1, H)) / H
    return np.array([float(c * M) for c in C]), float(H2)


def _k_forms(p, r1, r2):
    """K by the reference's rounding (h^2 = mean of |dr|^2 - t^2, two atan2) and by the
    kernels' accurate form (h^2 = |dr1 x e|^2, t2 = t1 - len, one atan2), in float64, and the
    shared rounded inputs."""
    d = r2 - r1
    length = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
    e = d / length
    d1, d2 = p - r1, p - r2
    c = np.array([d1[1] * e[2] - d1[2] * e[1], d1[2] * e[0] - d1[0] * e[2], d1[0] * e[1] - d1[1] * e[0]])
    t1 = e[0] * d1[0] + e[1] * d1[1] + e[2] * d1[2]
    t2 = e[0] * d2[0] + e[1] * d2[1] + e[2] * d2[2]
    h2r = ((d1 @ d1 - t1 * t1) + (d2 @ d2 - t2 * t2)) / 2.0
    hr = np.sqrt(h2r)
    k_ref = c * ((np.arctan2(t2, hr) - np.arctan2(t1, hr)) / hr)
    h2 = c @ c
    h = np.sqrt(h2)
    k_acc = c * (np.arctan2(-length * h, h2 + t1 * (t1 - length)) / h)
    return k_ref, k_acc, h2r, h2, (d1, e, length)


def test_cross_product_distance_form():
    """h^2 = |dr1 x e|^2 against the reference's ((|dr1|^2 - t1^2) + (|dr2|^2 - t2^2)) / 2.
    Points beside a segment and close to its line, in the far field, and near the line's
    extension.  Against 50-digit arithmetic on the rounded inputs both forms share (dr1, e,
    len): the cross product (which K's direction needs in both forms) carries an error of
    ~eps |dr1| / h relative, so the accurate form's h^2 is within a few eps |dr1| / h, while
    the reference form's difference loses ~eps |dr|^2 / h^2; the accurate form's K is within
    the cross product's own error of the exact K; and the two forms' K differ by no more than
    the reference's rounding of h^2 allows."""
    import mpmath as mp

    mp.mp.dps = 50
    rng = np.random.default_rng(11)
    eps = np.finfo(float).eps
    worst_h2_acc, worst_h2_ref = 0.0, 0.0
    for k in range(300):
        r1 = rng.normal(size=3)
        d = rng.normal(size=3)
        r2 = r1 + d
        e = d / np.linalg.norm(d)
        nrm = np.cross(e, rng.normal(size=3))
        nrm /= np.linalg.norm(nrm)
        kind = k % 3
        if kind == 0:  # beside the segment, close to its line
            p = r1 + rng.uniform(0.1, 0.9) * d + nrm * 10.0 ** rng.uniform(-5, -2)
        elif kind == 1:  # far field
            p = r1 + rng.normal(size=3) * 10.0 ** rng.uniform(0, 2)
        else:  # near the line's extension, far along it
            p = r1 + d * rng.uniform(3, 30) + nrm * 10.0 ** rng.uniform(-3, 0)
        k_ref, k_acc, h2r, h2, inputs = _k_forms(p, r1, r2)
        exact, h2x = _exact_k(*inputs)
        scale = np.max(np.abs(exact))
        ratio = np.linalg.norm(inputs[0]) / np.sqrt(h2x)  # |dr1| / h
        dr2 = max((p - r1) @ (p - r1), (p - r2) @ (p - r2))
        err_acc = abs(h2 - h2x) / h2x
        err_ref = abs(h2r - h2x) / h2x
        assert err_acc <= 8 * eps * (ratio + 1), k
        assert err_ref <= 8 * eps * dr2 / h2x + 8 * eps * (ratio + 1), k
        assert np.max(np.abs(k_acc - exact)) / scale <= 8 * eps * (ratio + 1) + 1e-14, k
        assert np.max(np.abs(k_acc - k_ref)) / scale <= 8 * eps * dr2 / h2x + 8 * eps * (ratio + 1) + 1e-14, k
        if kind == 0:
            worst_h2_acc, worst_h2_ref = max(worst_h2_acc, err_acc), max(worst_h2_ref, err_ref)
    assert worst_h2_ref > 100 * worst_h2_acc  # close to the line the accurate form is far better


def test_refined_rsqrt():
    """pvl_rsqrt: one third-order step y (1 + e/2 + 3 e^2/8), e = 1 - x y^2, from estimates
    off by up to 2^-22 relative (twice the hardware's stated bound) lands within 2 ulp of
    1/sqrt(x) (50-digit reference), over 30 decades."""
    import mpmath as mp

    mp.mp.dps = 40
    rng = np.random.default_rng(5)
    x = 10.0 ** rng.uniform(-15, 15, 3000)
    worst = 0.0
    for xi in x:
        exact = 1 / mp.sqrt(mp.mpf(float(xi)))
        y = float(exact) * (1.0 + rng.uniform(-1, 1) * 2.0 ** -22)
        e = 1.0 - (xi * y) * y  # the kernel fuses this into one FMA; a plain product is worse
        r = (y * e) * (0.375 * e + 0.5) + y
        worst = max(worst, float(abs((mp.mpf(r) - exact) / exact)))
    assert worst < 2 * 2.0 ** -52
