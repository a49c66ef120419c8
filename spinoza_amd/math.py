"""Numeric constants and helpers (spinoza/src/math.rs:5-37)."""
import math as _m

PI = _m.pi                      # math.rs:8
SQRT_ONE_HALF = 0.7071067811865476   # FRAC_1_SQRT_2, math.rs:5
Float = float                   # f64 (feature "double")


class Amplitude:
    """One complex amplitude (math.rs:18-24)."""
    __slots__ = ("re", "im")

    def __init__(self, re=0.0, im=0.0):
        self.re = float(re)
        self.im = float(im)

    def __repr__(self):
        return f"Amplitude(re={self.re}, im={self.im})"


def modulus(z_re, z_im):
    """|z| = sqrt(fma(re, re, im*im)) (math.rs:28-31)."""
    return _m.sqrt(_m.fma(z_re, z_re, z_im * z_im)) if hasattr(_m, "fma") else \
        _m.sqrt(z_re * z_re + z_im * z_im)


def pow2f(n):
    """2^n as a float (math.rs:33-37)."""
    return float(2.0 ** int(n))
