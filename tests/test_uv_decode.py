"""The uv decode of the level kernels evaluates the single-precision quotient
k / 65535.0f (CatmullClark.h:374-382, k a 16-bit code) as

    q = RN(k * r),  r = RN(1 / 65535)
    e = RN(k - q * 65535)        (fused multiply-add)
    result = RN(q + e * r)       (fused multiply-add)

with RN the round-to-nearest-even to binary32.  This test proves with exact
rational arithmetic that the result equals the correctly rounded quotient
for all 65536 codes (the three rounding steps are evaluated exactly, the
same way the hardware's fused multiply-add does).  The plain product
RN(k * r) differs for some codes, which the test also shows, so it would
catch a dropped correction step.
"""
from fractions import Fraction

import numpy as np


def rn32(x: Fraction) -> Fraction:
    """Round an exact rational to the nearest binary32 (ties to even)."""
    if x == 0:
        return Fraction(0)
    guess = np.float32(float(x))
    cands = {guess, np.nextafter(guess, np.float32(np.inf)), np.nextafter(guess, np.float32(-np.inf))}
    best = None
    for c in cands:
        err = abs(Fraction(float(c)) - x)
        key = (err, int(np.array(c, np.float32).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return Fraction(float(best[1]))


def test_newton_quotient_is_correctly_rounded_for_every_code():
    r = rn32(Fraction(1, 65535))
    plain_differs = 0
    for k in range(65536):
        x = Fraction(k)
        want = rn32(Fraction(k, 65535))
        q = rn32(x * r)
        e = rn32(x - q * 65535)
        got = rn32(q + e * r)
        assert got == want, k
        plain_differs += q != want
    assert plain_differs > 0


def test_unit_encode_truncation_matches_64_bit():
    # products of [0, 1] coordinates with 65535 lie in [0, 65535]; there the
    # 32-bit and 64-bit truncating conversions agree (including (-1, 0))
    xs = np.float32([0.0, 1e-30, 0.49999997, 0.5, 1.0, 65534.996, 65535.0, -0.5, -1e-30])
    for x in xs:
        assert int(np.int64(np.trunc(np.float64(x)))) & 0xFFFFFFFF == int(np.int32(np.trunc(np.float64(x)))) & 0xFFFFFFFF
