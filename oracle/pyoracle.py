"""ctypes binding of the C oracle (oracle/spz_oracle.c).

Test infrastructure only (see the package docstring).  Every function works on
float64 numpy arrays holding the split real/imaginary parts of a state, the
layout of the reference's ``State`` (spinoza/src/core.rs:18-51).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

H, X, Y, Z, P, RX, RY, RZ, U, SWAP, BITFLIP, M, UNITARY = range(13)
KIND = {"H": H, "X": X, "Y": Y, "Z": Z, "P": P, "RX": RX, "RY": RY,
        "RZ": RZ, "U": U, "SWAP": SWAP, "BITFLIP": BITFLIP, "M": M,
        "UNITARY": UNITARY}

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        up = ctypes.POINTER(ctypes.c_uint64)
        i, d, u64 = ctypes.c_int, ctypes.c_double, ctypes.c_uint64
        L.orc_apply.argtypes = [dp, dp, i, i, d, d, d, i, i, d]
        L.orc_swap.argtypes = [dp, dp, i, i, i]
        L.orc_c_apply.argtypes = [dp, dp, i, i, d, d, d, i, i]
        L.orc_mc_apply.argtypes = [dp, dp, i, i, d, d, d, u64, u64, i, i]
        L.orc_transform_u.argtypes = [dp, dp, i, i, dp, dp, i, i]
        L.orc_gate_matrix.argtypes = [i, d, d, d, dp]
        L.orc_norm.argtypes = [dp, dp, i]
        L.orc_norm.restype = d
        L.orc_prob0.argtypes = [dp, dp, i, i]
        L.orc_prob0.restype = d
        L.orc_qubit_expectation_value.argtypes = [dp, dp, i, i]
        L.orc_qubit_expectation_value.restype = d
        L.orc_xyz_expectation_value.argtypes = [dp, dp, i, ctypes.c_char, i, dp]
        L.orc_expect_zz.argtypes = [dp, dp, i, i, i]
        L.orc_expect_zz.restype = d
        L.orc_dot_re.argtypes = [dp, dp, dp, dp, i]
        L.orc_dot_re.restype = d
        L.orc_ising_apply.argtypes = [dp, dp, i, d, dp, dp]
        L.orc_ising_energy.argtypes = [dp, dp, i, d]
        L.orc_ising_energy.restype = d
        L.orc_measure_qubit.argtypes = [dp, dp, i, i, i, i, d, dp]
        L.orc_sample.argtypes = [dp, dp, i, dp, u64, up]
        L.orc_uniforms.argtypes = [u64, u64, dp]
        L.orc_gaussian_state.argtypes = [u64, i, dp, dp]
        L.orc_random_state.argtypes = [u64, i, dp, dp]
        L.orc_set_threads.argtypes = [i]
        L.orc_get_threads.restype = i
        L.orc_execute.argtypes = [dp, dp, i, ctypes.c_void_p, u64, up]
        L.orc_execute.restype = ctypes.c_int64
        L.orc_rayon_apply.argtypes = [dp, dp, i, i, d, d, d, i, u64, u64]
        L.orc_rayon_execute.argtypes = [dp, dp, i, ctypes.c_void_p, u64]
        L.orc_rayon_execute.restype = ctypes.c_int64
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleState:
    """Split re/im float64 arrays, |0...0> by default (core.rs:32-42)."""

    def __init__(self, n, re=None, im=None):
        self.n = n
        N = 1 << n
        if re is None:
            self.re = np.zeros(N)
            self.im = np.zeros(N)
            self.re[0] = 1.0
        else:
            self.re = np.ascontiguousarray(re, dtype=np.float64).copy()
            self.im = np.ascontiguousarray(im, dtype=np.float64).copy()

    def copy(self):
        return OracleState(self.n, self.re, self.im)

    def apply(self, kind, target, theta=0.0, phi=0.0, lam=0.0, t1=0, eps=1.0):
        rc = lib().orc_apply(_p(self.re), _p(self.im), self.n, KIND.get(kind, kind),
                             theta, phi, lam, target, t1, eps)
        if rc:
            raise ValueError(f"oracle apply failed: {kind} t={target}")

    def swap(self, t0, t1):
        if lib().orc_swap(_p(self.re), _p(self.im), self.n, t0, t1):
            raise ValueError("oracle swap failed")

    def c_apply(self, kind, control, target, theta=0.0, phi=0.0, lam=0.0):
        rc = lib().orc_c_apply(_p(self.re), _p(self.im), self.n, KIND.get(kind, kind),
                               theta, phi, lam, control, target)
        if rc:
            raise ValueError("oracle c_apply failed")

    def mc_apply(self, kind, ones, zeros, target, theta=0.0, phi=0.0, lam=0.0, ry_as_rx=False):
        rc = lib().orc_mc_apply(_p(self.re), _p(self.im), self.n, KIND.get(kind, kind),
                                theta, phi, lam, ones, zeros, target, int(ry_as_rx))
        if rc:
            raise ValueError("oracle mc_apply failed")

    def transform_u(self, u, t, control=-1):
        u = np.asarray(u, dtype=np.complex128)
        m = int(round(np.log2(u.shape[0])))
        ur = np.ascontiguousarray(u.real)
        ui = np.ascontiguousarray(u.imag)
        rc = lib().orc_transform_u(_p(self.re), _p(self.im), self.n, m, _p(ur), _p(ui), t, control)
        if rc:
            raise ValueError("oracle transform_u failed")

    def norm(self):
        return lib().orc_norm(_p(self.re), _p(self.im), self.n)

    def prob0(self, t):
        return lib().orc_prob0(_p(self.re), _p(self.im), self.n, t)

    def qubit_expectation_value(self, t):
        return lib().orc_qubit_expectation_value(_p(self.re), _p(self.im), self.n, t)

    def xyz_expectation_value(self, obs, targets):
        out = []
        for t in targets:
            v = ctypes.c_double()
            if lib().orc_xyz_expectation_value(_p(self.re), _p(self.im), self.n,
                                               obs.encode(), t, ctypes.byref(v)):
                raise ValueError("bad observable or target")
            out.append(v.value)
        return out

    def expect_zz(self, q0, q1):
        return lib().orc_expect_zz(_p(self.re), _p(self.im), self.n, q0, q1)

    def ising_energy(self, hx=1.0):
        return lib().orc_ising_energy(_p(self.re), _p(self.im), self.n, hx)

    def ising_apply(self, hx=1.0):
        lr = np.zeros_like(self.re)
        li = np.zeros_like(self.im)
        lib().orc_ising_apply(_p(self.re), _p(self.im), self.n, hx, _p(lr), _p(li))
        return lr, li

    def measure_qubit(self, target, reset=True, forced=-1, u=0.5):
        p0 = ctypes.c_double()
        v = lib().orc_measure_qubit(_p(self.re), _p(self.im), self.n, target, int(reset),
                                    forced, u, ctypes.byref(p0))
        if v < 0:
            raise ValueError("measurement of a zero-probability outcome")
        return v, p0.value

    def sample(self, uniforms):
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.zeros(len(u), dtype=np.uint64)
        rc = lib().orc_sample(_p(self.re), _p(self.im), self.n, _p(u), len(u),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
        if rc:
            raise ValueError("sampling a zero state")
        return out


def uniforms(seed, count):
    out = np.zeros(count)
    lib().orc_uniforms(seed, count, _p(out))
    return out


def gaussian_state(seed, n):
    s = OracleState(n)
    lib().orc_gaussian_state(seed, n, _p(s.re), _p(s.im))
    return s


def random_state(seed, n):
    s = OracleState(n)
    lib().orc_random_state(seed, n, _p(s.re), _p(s.im))
    return s


def set_threads(t):
    lib().orc_set_threads(int(t))


class _OrcOp(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("target", ctypes.c_int32),
                ("theta", ctypes.c_double), ("phi", ctypes.c_double), ("lam", ctypes.c_double),
                ("t1", ctypes.c_int32), ("classical", ctypes.c_int32),
                ("ones", ctypes.c_uint64), ("zeros", ctypes.c_uint64),
                ("u", ctypes.c_double), ("m", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("u_re", ctypes.c_void_p), ("u_im", ctypes.c_void_p)]


def _encode(ops):
    arr = (_OrcOp * max(1, len(ops)))()
    keep = []
    for i, o in enumerate(ops):
        r = arr[i]
        r.kind = KIND.get(o["kind"], o["kind"]) if isinstance(o["kind"], str) else o["kind"]
        r.target = o.get("target", 0)
        r.theta = o.get("theta", 0.0)
        r.phi = o.get("phi", 0.0)
        r.lam = o.get("lam", 0.0)
        r.t1 = o.get("t1", 0)
        r.classical = int(o.get("classical", 0))
        r.ones = o.get("ones", 0)
        r.zeros = o.get("zeros", 0)
        r.u = o.get("u", 1.0)
        if "matrix" in o:
            u = np.asarray(o["matrix"], dtype=np.complex128)
            ur = np.ascontiguousarray(u.real)
            ui = np.ascontiguousarray(u.imag)
            keep += [ur, ui]
            r.m = int(round(np.log2(u.shape[0])))
            r.u_re = ur.ctypes.data
            r.u_im = ui.ctypes.data
    return arr, keep


def rayon_execute(state, ops):
    """The CPU baseline: an op list of uncontrolled and single-controlled
    gates through the reference's per-kind parallel loop shapes
    (orc_rayon_apply).  Results equal execute()'s bit for bit, except CZ,
    which runs as c_apply(P(pi)) as in the reference."""
    arr, keep = _encode(ops)
    rc = lib().orc_rayon_execute(_p(state.re), _p(state.im), state.n, ctypes.addressof(arr),
                                 len(ops))
    if rc:
        raise ValueError(f"CPU baseline does not cover op {-rc - 1}")


def execute(state, ops, tracker=None):
    """Run an op list (dicts with kind/target/theta/phi/lam/t1/ones/zeros/u/
    classical) through the oracle's QuantumCircuit::execute restatement."""
    arr, keep = _encode(ops)
    tr = np.zeros(2, dtype=np.uint64) if tracker is None else tracker
    rc = lib().orc_execute(_p(state.re), _p(state.im), state.n, ctypes.addressof(arr), len(ops),
                           tr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    if rc:
        raise ValueError(f"oracle execute failed at op {-rc - 1}")
    return tr
