"""Single-qubit measurement (spinoza/src/measurement.rs:12-92)."""
import ctypes

from . import _native as N
from .rng import default_rng


def measure_qubit(state, target, reset, v=None, u=None):
    """measure_qubit(&mut state, target, reset, v) (measurement.rs:12-92).

    ``v`` forces the outcome (the reference's ``Some(v)``).  Otherwise the
    outcome is 1 iff ``u < 1 - prob0``, with ``u`` drawn from the seeded
    process generator unless given.  The kept half is scaled by
    1/sqrt(p) and the other half zeroed; outcome 1 with ``reset`` applies X.
    Forcing an outcome of probability 0 raises SpinozaError (the reference
    would divide by zero)."""
    if v is not None and v not in (0, 1):
        raise ValueError("v must be 0 or 1 (measurement.rs:32)")
    if u is None:
        u = default_rng().uniform() if v is None else 0.0
    out = ctypes.c_uint8()
    N.check(N.lib().spz_measure_qubit(state._h, int(target), int(bool(reset)),
                                      -1 if v is None else int(v), float(u),
                                      ctypes.byref(out)))
    return int(out.value)
