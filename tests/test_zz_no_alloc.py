"""spz_ising_energy_noalloc (the entry point that allocates nothing) against
spz_ising_energy, through the C ABI: bit-identical energies on two calls in a
row, the second on a smaller state (the static scratch keeps the first call's
partials, which the second must not read), for FP64 and FP32 storage, block
counts of 1 and many; a state past the 2^30-amplitude limit is rejected and
the library still works afterwards.  Collected last (its file name) so that a
failure here hides no other test."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def energies(spz, s, hx):
    from spinoza_amd import _native as N
    a, b = ctypes.c_double(), ctypes.c_double()
    N.check(N.lib().spz_ising_energy(s._h, ctypes.c_double(hx), ctypes.byref(a)))
    N.check(N.lib().spz_ising_energy_noalloc(s._h, ctypes.c_double(hx), ctypes.byref(b)))
    return a.value, b.value


def prepared(spz, n, precision, seed):
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import random_circuit
    s = spz.State(n, precision)
    run_ops(s, random_circuit(n, 2, seed))
    return s


def test_noalloc_energy_matches_allocating_entry_point():
    import spinoza_amd as spz
    from spinoza_amd import _native as N
    print("case noalloc ising energy", flush=True)
    for precision in ("f64", "f32"):
        for n, hx in [(22, 0.7), (12, 1.3), (9, 0.4), (20, 1.0)]:
            s = prepared(spz, n, precision, seed=n)
            ref, got = energies(spz, s, hx)
            assert got == ref, (precision, n, got, ref)
            del s
    big = spz.State(31, "f32")
    out = ctypes.c_double()
    rc = N.lib().spz_ising_energy_noalloc(big._h, ctypes.c_double(1.0), ctypes.byref(out))
    assert rc == N.ERR_ARG and b"2^30" in N.lib().spz_last_error()
    del big
    s = prepared(spz, 14, "f64", seed=3)
    ref, got = energies(spz, s, 1.0)
    assert got == ref
