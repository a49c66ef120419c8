"""Quantum logic gates and their application to a GPU-resident State.

Mirrors spinoza/src/gates.rs: the ``Gate`` enum (:45-74), ``inverse``
(:78-92), ``to_matrix`` (:95-190) and the functional entry points ``apply``
(:215-232), ``c_apply`` (:257-269), ``cc_apply`` (:272-277), ``mc_apply``
(:290-320), ``transform_u`` (:1427-1452) and ``c_transform_u`` (:1455-1487).
The pair arithmetic runs in the gfx950 library; errors that the reference
reports by panicking raise ``SpinozaError`` (or ``ValueError`` for the
reference's ``unimplemented!``/``todo!`` cases).
"""
import math

import numpy as np

from . import _native as N
from .math import SQRT_ONE_HALF
from .rng import default_rng

_NAMES = {N.H: "H", N.X: "X", N.Y: "Y", N.Z: "Z", N.P: "P", N.RX: "RX", N.RY: "RY",
          N.RZ: "RZ", N.U: "U", N.SWAP: "SWAP", N.BITFLIP: "BitFlipNoise", N.M: "M",
          N.UNITARY: "Unitary"}


class Gate:
    """One gate.  Parameterless gates are class attributes (``Gate.H``);
    parameterised ones are constructors (``Gate.RX(theta)``,
    ``Gate.U(theta, phi, lam)``, ``Gate.SWAP(t0, t1)``,
    ``Gate.BitFlipNoise(prob)``, ``Gate.Unitary(u)``)."""

    __slots__ = ("kind", "theta", "phi", "lam", "t0", "t1", "unitary")

    def __init__(self, kind, theta=0.0, phi=0.0, lam=0.0, t0=0, t1=0, unitary=None):
        self.kind = kind
        self.theta = float(theta)
        self.phi = float(phi)
        self.lam = float(lam)
        self.t0 = int(t0)
        self.t1 = int(t1)
        self.unitary = unitary

    # -- constructors for the parameterised variants --
    @staticmethod
    def P(theta):
        return Gate(N.P, theta)

    @staticmethod
    def RX(theta):
        return Gate(N.RX, theta)

    @staticmethod
    def RY(theta):
        return Gate(N.RY, theta)

    @staticmethod
    def RZ(theta):
        return Gate(N.RZ, theta)

    @staticmethod
    def U(theta, phi, lam):
        return Gate(N.U, theta, phi, lam)

    @staticmethod
    def SWAP(t0, t1):
        return Gate(N.SWAP, t0=t0, t1=t1)

    @staticmethod
    def BitFlipNoise(prob):
        return Gate(N.BITFLIP, prob)

    @staticmethod
    def Unitary(u):
        return Gate(N.UNITARY, unitary=u)

    @property
    def name(self):
        return _NAMES[self.kind]

    def __repr__(self):
        if self.kind in (N.P, N.RX, N.RY, N.RZ, N.BITFLIP):
            return f"Gate.{self.name}({self.theta})"
        if self.kind == N.U:
            return f"Gate.U({self.theta}, {self.phi}, {self.lam})"
        if self.kind == N.SWAP:
            return f"Gate.SWAP({self.t0}, {self.t1})"
        return f"Gate.{self.name}"

    def __eq__(self, other):
        return (isinstance(other, Gate) and self.kind == other.kind
                and self.theta == other.theta and self.phi == other.phi
                and self.lam == other.lam and self.t0 == other.t0 and self.t1 == other.t1
                and self.unitary is other.unitary)

    def __hash__(self):
        return hash((self.kind, self.theta, self.phi, self.lam, self.t0, self.t1))

    def clone(self):
        return Gate(self.kind, self.theta, self.phi, self.lam, self.t0, self.t1,
                    self.unitary.clone() if self.unitary is not None else None)

    def inverse(self):
        """Gate::inverse (gates.rs:78-92)."""
        k = self.kind
        if k in (N.H, N.X, N.Y, N.Z, N.SWAP):
            return self
        if k in (N.P, N.RX, N.RY, N.RZ):
            return Gate(k, -self.theta)
        if k == N.U:
            return Gate(N.U, -self.theta, -self.lam, -self.phi)
        if k == N.UNITARY:
            u = self.unitary.clone()
            u.conj_t()
            return Gate(N.UNITARY, unitary=u)
        raise ValueError(f"{self.name} has no inverse (gates.rs:86 unimplemented!)")

    def to_matrix(self):
        """Gate::to_matrix (gates.rs:95-190) as a 2x2 complex numpy array."""
        k = self.kind
        if k == N.H:
            h = SQRT_ONE_HALF
            return np.array([[h, h], [h, -h]], dtype=complex)
        if k == N.X:
            return np.array([[0, 1], [1, 0]], dtype=complex)
        if k == N.Y:
            return np.array([[0, -1j], [1j, 0]], dtype=complex)
        if k == N.Z:
            return np.array([[1, 0], [0, -1]], dtype=complex)
        th = self.theta
        if k == N.P:
            return np.array([[1, 0], [0, complex(math.cos(th), math.sin(th))]], dtype=complex)
        t = th / 2.0
        c, s = math.cos(t), math.sin(t)
        if k == N.RX:
            return np.array([[c, complex(0, -s)], [complex(0, -s), c]], dtype=complex)
        if k == N.RY:
            return np.array([[c, -s], [s, c]], dtype=complex)
        if k == N.RZ:
            return np.array([[complex(c, -s), 0], [0, complex(c, s)]], dtype=complex)
        if k == N.U:
            ph, la = self.phi, self.lam
            return np.array([
                [complex(c, 0.0), complex(-math.cos(la) * s, -math.sin(la) * s)],
                [complex(math.cos(ph) * s, math.sin(ph) * s),
                 complex(math.cos(ph + la) * c, math.sin(ph + la) * c)]], dtype=complex)
        raise ValueError(f"{self.name} has no 2x2 matrix (gates.rs:188 unimplemented!)")

    def _ctype(self):
        g = N.Gate_t()
        g.kind = self.kind
        g.t1 = self.t1
        g.theta = self.theta
        g.phi = self.phi
        g.lam = self.lam
        return g


for _k, _n in ((N.H, "H"), (N.X, "X"), (N.Y, "Y"), (N.Z, "Z"), (N.M, "M")):
    setattr(Gate, _n, Gate(_k))


def apply(gate, state, target, eps=None):
    """gates::apply (gates.rs:215-232).  For BitFlipNoise ``eps`` is the
    uniform draw (default: the seeded process generator); SWAP ignores
    ``target`` as in the reference."""
    if gate.kind in (N.M, N.UNITARY):
        raise ValueError("apply: M and Unitary are not supported (gates.rs:230 unimplemented!)")
    if gate.kind == N.SWAP:
        g = gate._ctype()
        N.check(N.lib().spz_apply(state._h, g, gate.t0, 1.0))
        return
    if gate.kind == N.BITFLIP and eps is None:
        eps = default_rng().uniform()
    N.check(N.lib().spz_apply(state._h, gate._ctype(), int(target),
                              1.0 if eps is None else float(eps)))


def c_apply(gate, state, control, target):
    """gates::c_apply (gates.rs:257-269)."""
    if gate.kind > N.U:
        raise ValueError(f"c_apply: {gate.name} unsupported (gates.rs:267 todo!)")
    N.check(N.lib().spz_c_apply(state._h, gate._ctype(), int(control), int(target)))


def cc_apply(gate, state, control0, control1, target):
    """gates::cc_apply (gates.rs:272-277); any single-target gate."""
    if gate.kind > N.U:
        raise ValueError(f"cc_apply: {gate.name} unsupported (gates.rs:275 todo!)")
    N.check(N.lib().spz_cc_apply(state._h, gate._ctype(), int(control0), int(control1),
                                 int(target)))


def _mask(qubits):
    m = 0
    for q in qubits:
        m |= 1 << int(q)
    return m


def mc_apply(gate, state, controls, zeros, target):
    """gates::mc_apply (gates.rs:290-320), reproducing the reference's mask
    construction: qubits listed in ``zeros`` are dropped from the control
    mask (gates.rs:300-306), so they are ignored rather than required to
    read 0.  Use ``mc_apply_masked`` for true 0-controls."""
    zs = set(zeros) if zeros is not None else set()
    ones = _mask(c for c in controls if c not in zs)
    mc_apply_masked(gate, state, ones, 0, target)


def mc_apply_masked(gate, state, ones, zeros, target):
    """Apply `gate` to `target` on the amplitudes whose qubits in the bit mask
    ``ones`` read 1 and whose qubits in ``zeros`` read 0."""
    if gate.kind > N.U:
        raise ValueError(f"mc_apply: {gate.name} unsupported (gates.rs:318 todo!)")
    N.check(N.lib().spz_mc_apply(state._h, gate._ctype(), int(ones), int(zeros), int(target)))


def _upload_block(state, u, t, control):
    re = np.ascontiguousarray(u.reals, dtype=np.float64)
    im = np.ascontiguousarray(u.imags, dtype=np.float64)
    if u.height != u.width:
        raise ValueError("unitary must be square")
    m = int(u.height).bit_length() - 1
    if (1 << m) != u.height:
        raise ValueError("unitary size must be a power of two")
    N.check(N.lib().spz_transform_u(state._h, m, N.dptr(re), N.dptr(im), int(t), int(control)))


def transform_u(state, u, t):
    """gates::transform_u (gates.rs:1427-1452): a 2^m x 2^m block on qubits
    [t, t+m) of the state (any t; the reference is correct only at t = 0 with
    a full-state matrix)."""
    _upload_block(state, u, t, -1)


def c_transform_u(state, u, c, t):
    """gates::c_transform_u (gates.rs:1455-1487); the control must lie
    outside the block."""
    _upload_block(state, u, t, c)
