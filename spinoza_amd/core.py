"""GPU-resident State, expectation values, IQFT and sampling.

Mirrors spinoza/src/core.rs: ``State`` (:18-51), ``Reservoir`` /
``reservoir_sampling`` (:65-129), ``iqft`` (:184-191),
``qubit_expectation_value`` (:198-219), ``xyz_expectation_value`` (:222-264).
"""
import ctypes

import numpy as np

from . import _native as N
from .math import PI, pow2f
from .rng import default_rng


class State:
    """2^n amplitudes in device memory as two arrays, reals and imags.

    ``State(n)`` is |0...0> (core.rs:32-42).  ``reals`` / ``imags`` return host
    copies (the reference exposes its Vecs directly); assigning a whole array
    to them uploads it.  ``precision`` is ``"f64"`` (default) or ``"f32"``.
    """

    def __init__(self, n, precision="f64"):
        n = int(n)
        if n <= 0:
            raise ValueError("State::new: n must be > 0 (core.rs:33)")
        N.load()
        prec = N.F64 if precision in ("f64", N.F64) else N.F32
        h = ctypes.c_void_p()
        N.check(N.lib().spz_state_create(n, prec, ctypes.byref(h)))
        self._h = h
        self.n = n
        self.precision = "f64" if prec == N.F64 else "f32"

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and N._lib is not None and not N.exiting():
            N._lib.spz_state_destroy(h)
        self._h = None

    def __len__(self):
        return 1 << self.n

    def len(self):
        return 1 << self.n

    # -- host mirror ----------------------------------------------------------
    def download(self, offset=0, count=None):
        count = (1 << self.n) - offset if count is None else count
        re = np.empty(count)
        im = np.empty(count)
        N.check(N.lib().spz_state_download(self._h, N.dptr(re), N.dptr(im), offset, count))
        return re, im

    def upload(self, re, im, offset=0):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        if re.shape != im.shape:
            raise ValueError("reals and imags must have the same length")
        N.check(N.lib().spz_state_upload(self._h, N.dptr(re), N.dptr(im), offset, len(re)))

    @property
    def reals(self):
        return self.download()[0]

    @reals.setter
    def reals(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        N.check(N.lib().spz_state_upload(self._h, N.dptr(v), None, 0, len(v)))

    @property
    def imags(self):
        return self.download()[1]

    @imags.setter
    def imags(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        N.check(N.lib().spz_state_upload(self._h, None, N.dptr(v), 0, len(v)))

    def __getitem__(self, i):
        re, im = self.download(int(i), 1)
        return float(re[0]), float(im[0])

    def to_numpy(self):
        re, im = self.download()
        return re + 1j * im

    @classmethod
    def from_arrays(cls, re, im, precision="f64"):
        re = np.asarray(re, dtype=np.float64)
        n = int(len(re)).bit_length() - 1
        if (1 << n) != len(re):
            raise ValueError("length must be a power of two")
        s = cls(n, precision)
        s.upload(re, im)
        return s

    def set_basis(self, index):
        N.check(N.lib().spz_state_set_basis(self._h, int(index)))

    def clone(self):
        s = State(self.n, self.precision)
        N.check(N.lib().spz_state_copy(s._h, self._h))
        return s

    def copy_from(self, other):
        N.check(N.lib().spz_state_copy(self._h, other._h))

    def device_ptrs(self):
        re, im = ctypes.c_void_p(), ctypes.c_void_p()
        N.check(N.lib().spz_state_device_ptrs(self._h, ctypes.byref(re), ctypes.byref(im)))
        return re.value, im.value

    def norm(self):
        out = ctypes.c_double()
        N.check(N.lib().spz_norm(self._h, ctypes.byref(out)))
        return out.value

    def __str__(self):
        re, im = self.download()
        return "".join(f"{a} + i{b}\n" for a, b in zip(re, im))


def iqft(state, targets):
    """core::iqft (core.rs:184-191)."""
    from .gates import Gate, apply, c_apply
    for j in range(len(targets) - 1, -1, -1):
        apply(Gate.H, state, targets[j])
        for k in range(j - 1, -1, -1):
            c_apply(Gate.P(-PI / pow2f(j - k)), state, targets[j], targets[k])


def prob0(state, target):
    out = ctypes.c_double()
    N.check(N.lib().spz_prob0(state._h, int(target), ctypes.byref(out)))
    return out.value


def qubit_expectation_value(state, target):
    """core::qubit_expectation_value (core.rs:198-219): 2*prob0 - 1."""
    out = ctypes.c_double()
    N.check(N.lib().spz_qubit_expectation_value(state._h, int(target), ctypes.byref(out)))
    return out.value


def xyz_expectation_value(observable, state, targets):
    """core::xyz_expectation_value (core.rs:222-264): Re <psi|O_t|psi> for
    each target, O in 'x', 'y', 'z'.  A bad observable raises ValueError (the
    reference panics, core.rs:223-225)."""
    if observable not in ("x", "y", "z"):
        raise ValueError(f"observable {observable} not supported")
    t = np.ascontiguousarray(targets, dtype=np.int32)
    out = np.zeros(len(t))
    N.check(N.lib().spz_xyz_expectation_value(
        state._h, observable.encode(), t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
        len(t), N.dptr(out)))
    return [float(v) for v in out]


def expect_zz(state, q0, q1):
    out = ctypes.c_double()
    N.check(N.lib().spz_expect_zz(state._h, int(q0), int(q1), ctypes.byref(out)))
    return out.value


def ising_energy(state, hx=1.0):
    """<psi|H|psi> for H = sum_{i<n-1} Z_i Z_{i+1} + hx sum_i X_i."""
    out = ctypes.c_double()
    N.check(N.lib().spz_ising_energy(state._h, float(hx), ctypes.byref(out)))
    return out.value


def inner(a, b):
    """<a|b>."""
    re, im = ctypes.c_double(), ctypes.c_double()
    N.check(N.lib().spz_inner(a._h, b._h, ctypes.byref(re), ctypes.byref(im)))
    return complex(re.value, im.value)


def sample(state, uniforms):
    """Indices drawn from |psi|^2 by the inverse-CDF rule, one per uniform."""
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint64)
    N.check(N.lib().spz_sample(state._h, N.dptr(u), len(u),
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
    return out


class Reservoir:
    """Sampler result with the reference's Reservoir interface
    (core.rs:65-121): ``entries`` holds k sampled indices."""

    def __init__(self, k):
        self.entries = np.zeros(int(k), dtype=np.uint64)
        self.w_s = 0.0

    def sampling(self, state, num_tests=None, rng=None):
        """Fill the k slots with indices distributed as |psi|^2.  The
        reference's weighted reservoir converges to that distribution as
        num_tests grows; the inverse-CDF sampler is exact, so num_tests is
        accepted for interface compatibility only."""
        rng = rng or default_rng()
        self.entries = sample(state, rng.uniforms(len(self.entries)))
        self.w_s = 1.0

    def get_outcome_count(self):
        """Histogram {index: count} (core.rs:114-121)."""
        vals, counts = np.unique(self.entries, return_counts=True)
        return {int(v): int(c) for v, c in zip(vals, counts)}


def reservoir_sampling(state, reservoir_size, num_tests=None, rng=None):
    """core::reservoir_sampling (core.rs:125-129)."""
    r = Reservoir(reservoir_size)
    r.sampling(state, num_tests, rng)
    return r
