"""The FP64 back substitution forms x / u_kk as q0 = x r, e = fma(-q0, u_kk, x),
q = fma(e, r, q0) with r = 1 / u_kk (a true, correctly rounded division done once per row,
off the dependency chain).  With a correctly rounded reciprocal this is Markstein's
correction and returns the correctly rounded quotient wherever the residual e does not
underflow (|x| >= 2^-969; solution entries are nowhere near that).  This CPU test replays the sequence
operation by operation (products and FMAs in exact rational arithmetic, one rounding each,
as the hardware does) and requires it to equal IEEE division, on random operands over a wide
exponent range and on quotients near rounding boundaries; the uncorrected product x r must
differ somewhere in the same sample, so the test would catch a dropped correction."""

import random
from fractions import Fraction

import numpy as np


def _rn(v: Fraction) -> float:
    return float(v)  # Fraction -> float rounds to nearest, ties to even


def _markstein(x: float, d: float) -> float:
    r = 1.0 / d
    q0 = x * r
    e = _rn(Fraction(x) - Fraction(d) * Fraction(q0))
    return _rn(Fraction(e) * Fraction(r) + Fraction(q0))


def _operands():
    rng = random.Random(7)
    for _ in range(6000):
        yield rng.uniform(-1, 1) * 2.0 ** rng.randint(-60, 60), rng.uniform(0.5, 1) * 2.0 ** rng.randint(-60, 60)
    # quotients within a few ulps of a rounding midpoint: x = d * (q + ulp(q) / 2) rounded
    for _ in range(4000):
        d = rng.uniform(1, 2) * 2.0 ** rng.randint(-20, 20)
        q = rng.uniform(1, 2)
        mid = Fraction(q) + Fraction(np.spacing(q)) / 2
        x = _rn(Fraction(d) * mid)
        for k in (-2, -1, 0, 1, 2):
            yield float(np.nextafter(x, np.inf) if k > 0 else np.nextafter(x, -np.inf) if k < 0 else x), d
    for d in (3.0, 7.0, 0.1, 1.0 + 2.0 ** -52, 2.0 - 2.0 ** -51, 1e-300, 1e300):
        for x in (1.0, -1.0, 1.0 + 2.0 ** -52, 0.3, 5e-310 if d < 1 else 1e-10):
            yield x, d


def test_markstein_quotient_is_correctly_rounded():
    """Exact wherever the residual e (~|x| 2^-53) stays normal, i.e. |x| >= 2^-969; below
    that, e loses bits to underflow and the quotient is within one ulp."""
    bad, tiny, uncorrected_differs = 0, 0, 0
    for x, d in _operands():
        want = x / d
        got = _markstein(x, d)
        if abs(x) >= 2.0 ** -969:
            bad += got != want
        else:
            tiny += 1
            assert abs(got - want) <= np.spacing(abs(want))
        uncorrected_differs += x * (1.0 / d) != want
    assert bad == 0
    assert tiny > 0
    assert uncorrected_differs > 0
