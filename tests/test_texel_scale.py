"""The search's texel value (u8 / 255) * 65536 (compression_pipeline.cpp:44),
which the kernels form as fma(u8, RN(1/255), 257 * u8) instead of dividing:
equal to the correctly rounded quotient for all 256 inputs, checked in exact
rational arithmetic (a float32 fused multiply-add rounds the exact
a * b + c once)."""
from fractions import Fraction

import numpy as np


def _rn32(fr):
    """Round-to-nearest-even of an exact rational to float32."""
    d = np.float32(float(fr))
    best = None
    for c in (np.nextafter(d, np.float32(-np.inf)), d, np.nextafter(d, np.float32(np.inf))):
        key = (abs(Fraction(float(c)) - fr), int(np.array([c], np.float32).view(np.int32)[0]) & 1)
        if best is None or key < best[0]:
            best = (key, c)
    return best[1]


def test_fused_form_equals_the_quotient_for_every_byte():
    inv255 = np.float32(1.0) / np.float32(255.0)  # RN(1/255), as the compiler folds 1.0f / 255.0f
    for u in range(256):
        quotient = (np.float32(u) / np.float32(255.0)) * np.float32(65536.0)  # the oracle's form
        assert quotient == _rn32(Fraction(u * 65536, 255))
        x257 = np.float32(u) * np.float32(257.0)
        assert Fraction(float(x257)) == u * 257  # exact
        fused = _rn32(Fraction(float(np.float32(u))) * Fraction(float(inv255)) + Fraction(float(x257)))
        assert fused == quotient, u


def test_plain_product_is_not_enough():
    """A single rounded multiply by RN(65536/255) is wrong for about half of
    the inputs, which is why the kernels keep the fused form."""
    c = np.float32(65536.0) / np.float32(255.0)
    u = np.arange(256, dtype=np.float32)
    assert int(np.sum(u * c != (u / np.float32(255.0)) * np.float32(65536.0))) > 100
