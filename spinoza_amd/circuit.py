"""Quantum circuits: registers, controls, transformations and the executor.

Mirrors spinoza/src/circuit.rs: ``QuantumRegister`` (:15-51), ``Controls``
(:55-109), ``QuantumTransformation`` (:112-120), ``QubitTracker``
(:122-164) and ``QuantumCircuit`` (:166-601).  ``execute`` hands the whole
op list to the gfx950 executor in one call (``spz_execute``), which is the
batching point where gates are grouped into fused tile sweeps.
"""
import ctypes

import numpy as np

from . import _native as N
from .core import State
from .gates import Gate
from .math import PI, pow2f
from .rng import default_rng


class QuantumRegister:
    """circuit.rs:15-51: the qubit indices of one register."""

    def __init__(self, size):
        size = int(size)
        if size <= 0:
            raise ValueError("QuantumRegister::new: size must be > 0 (circuit.rs:29)")
        self.qubits = list(range(size))

    def __getitem__(self, i):
        return self.qubits[i]

    def __len__(self):
        return len(self.qubits)

    def len(self):
        return len(self.qubits)

    def update_shift(self, shift):
        self.qubits = [q + shift for q in self.qubits]

    def get_shift(self):
        return self.qubits[0]


class Controls:
    """circuit.rs:55-109: None / Single / Ones / Mixed."""

    NONE, SINGLE, ONES, MIXED = range(4)

    def __init__(self, kind=0, controls=(), zeros=()):
        self.kind = kind
        self.controls = list(controls)
        self.zeros = set(zeros)

    @staticmethod
    def none():
        return Controls(Controls.NONE)

    @staticmethod
    def single(c):
        return Controls(Controls.SINGLE, [c])

    @staticmethod
    def ones(cs):
        return Controls(Controls.ONES, cs)

    @staticmethod
    def mixed(cs, zeros):
        return Controls(Controls.MIXED, cs, zeros)

    @staticmethod
    def from_lists(controls, zeros=None):
        """Controls::from (circuit.rs:72-86)."""
        if zeros is not None:
            return Controls.mixed(controls, zeros)
        if len(controls) == 0:
            return Controls.none()
        if len(controls) == 1:
            return Controls.single(controls[0])
        return Controls.ones(controls)

    def unpack(self):
        return list(self.controls), set(self.zeros)

    def new_with_control(self, control, shift):
        """circuit.rs:97-108."""
        controls, zeros = self.unpack()
        controls = [c + shift for c in controls] + [control]
        if not zeros:
            return Controls.from_lists(controls, None)
        return Controls.from_lists(controls, {z + shift for z in zeros})

    def clone(self):
        return Controls(self.kind, self.controls, self.zeros)


class QuantumTransformation:
    """circuit.rs:112-120."""

    def __init__(self, gate, target, controls=None):
        self.gate = gate
        self.target = int(target)
        self.controls = controls if controls is not None else Controls.none()


class QubitTracker:
    """circuit.rs:122-164 (backed by the C ABI's spz_tracker)."""

    def __init__(self):
        self.t = N.Tracker_t(0, 0)

    def is_qubit_measured(self, q):
        return bool((self.t.measured >> q) & 1)

    def get_qubit_measured_val(self, q):
        if not self.is_qubit_measured(q):
            return None
        return int((self.t.values >> q) & 1)


def _mask(qs):
    m = 0
    for q in qs:
        m |= 1 << int(q)
    return m


class QuantumCircuit:
    """circuit.rs:166-601.  ``QuantumCircuit(qr0, qr1, ...)`` (or a list of
    registers) shifts each register after the previous ones, as
    ``QuantumCircuit::new`` does, and allocates the state on the GPU."""

    def __init__(self, *registers, state=None, precision="f64"):
        if len(registers) == 1 and isinstance(registers[0], (list, tuple)):
            registers = tuple(registers[0])
        bits = 0
        sizes = []
        for r in registers:
            r.update_shift(bits)
            sizes.append(len(r))
            bits += len(r)
        self.transformations = []
        self.quantum_registers_info = sizes
        self.qubit_tracker = QubitTracker()
        if state is not None:
            self.state = state
        else:
            if bits == 0:
                raise ValueError("a circuit needs at least one qubit")
            self.state = State(bits, precision)
        self.exec_mode = N.EXEC_AUTO
        self.rng = default_rng()
        self.last_outcomes = None

    @classmethod
    def with_state(cls, state):
        """A circuit over an existing state (the reference's tests build
        QuantumCircuit { state, .. } directly)."""
        return cls(state=state)

    @property
    def num_qubits(self):
        return self.state.n

    def get_statevector(self):
        return self.state

    def inverse(self):
        """circuit.rs:206-211."""
        self.transformations.reverse()
        for qt in self.transformations:
            qt.gate = qt.gate.clone().inverse()

    def add(self, transformation):
        self.transformations.append(transformation)

    def _add(self, gate, target, controls=None):
        self.add(QuantumTransformation(gate, target, controls))

    # -- builders (circuit.rs:215-445) ---------------------------------------
    def measure(self, target):
        self._add(Gate.M, target)

    def swap(self, t0, t1):
        self._add(Gate.SWAP(t0, t1), 0)

    def x(self, target):
        self._add(Gate.X, target)

    def y(self, target):
        self._add(Gate.Y, target)

    def z(self, target):
        self._add(Gate.Z, target)

    def h(self, target):
        self._add(Gate.H, target)

    def p(self, angle, target):
        self._add(Gate.P(angle), target)

    def rx(self, angle, target):
        self._add(Gate.RX(angle), target)

    def ry(self, angle, target):
        self._add(Gate.RY(angle), target)

    def rz(self, angle, target):
        self._add(Gate.RZ(angle), target)

    def u(self, theta, phi, lam, target):
        self._add(Gate.U(theta, phi, lam), target)

    def cx(self, control, target):
        self._add(Gate.X, target, Controls.single(control))

    def ccx(self, control1, control2, target):
        self._add(Gate.X, target, Controls.ones([control1, control2]))

    def ch(self, control, target):
        self._add(Gate.H, target, Controls.single(control))

    def cy(self, control, target):
        self._add(Gate.Y, target, Controls.single(control))

    def cz(self, control, target):
        """Controlled Z (no builder in the reference, whose c_apply lacks a Z
        arm; it is supported here)."""
        self._add(Gate.Z, target, Controls.single(control))

    def cp(self, angle, control, target):
        self._add(Gate.P(angle), target, Controls.single(control))

    def crx(self, angle, control, target):
        self._add(Gate.RX(angle), target, Controls.single(control))

    def cry(self, angle, control, target):
        self._add(Gate.RY(angle), target, Controls.single(control))

    def crz(self, angle, control, target):
        self._add(Gate.RZ(angle), target, Controls.single(control))

    def cu(self, theta, phi, lam, control, target):
        self._add(Gate.U(theta, phi, lam), target, Controls.single(control))

    def bit_flip_noise(self, prob, target):
        self._add(Gate.BitFlipNoise(prob), target)

    def iqft(self, targets):
        """circuit.rs:438-445."""
        for j in range(len(targets) - 1, -1, -1):
            self.h(targets[j])
            for k in range(j - 1, -1, -1):
                self.cp(-PI / pow2f(j - k), targets[j], targets[k])

    def unitary(self, u, target):
        self._add(Gate.Unitary(u), target)

    def append_u(self, u, qr):
        if u.height != u.width or u.height != (1 << len(qr)):
            raise ValueError("unitary does not match the register (circuit.rs:524-525)")
        self.unitary(u, qr.get_shift())

    def c_unitary(self, u, c, t):
        self._add(Gate.Unitary(u), t, Controls.single(c))

    def c_append_u(self, u, c, qr):
        if u.height != u.width or u.height != (1 << len(qr)):
            raise ValueError("unitary does not match the register (circuit.rs:539-540)")
        self.c_unitary(u, c, qr.get_shift())

    def append(self, circuit, reg):
        """circuit.rs:448-460."""
        if len(reg) != sum(circuit.quantum_registers_info):
            raise ValueError("register size does not match the circuit (circuit.rs:450)")
        for tr in circuit.transformations:
            self._add(tr.gate.clone(), reg.get_shift() + tr.target, tr.controls.clone())

    def c_append(self, circuit, c, reg):
        """circuit.rs:463-477."""
        if reg.get_shift() <= c < reg.get_shift() + len(reg):
            raise ValueError("control lies inside the register (circuit.rs:464-468)")
        for tr in circuit.transformations:
            self._add(tr.gate.clone(), reg.get_shift() + tr.target,
                      tr.controls.new_with_control(c, reg.get_shift()))

    def mc_append(self, circuit, controls, reg):
        """circuit.rs:480-511 (one copy of the circuit per control, as in the
        reference)."""
        if len(set(controls)) != len(controls):
            raise ValueError("repeated control (circuit.rs:488)")
        lo, hi = reg.get_shift(), reg.get_shift() + len(reg)
        for c in controls:
            if lo <= c < hi:
                raise ValueError(f"control {c} should not be in: Range(start: {lo} end: {hi})")
        for control in controls:
            for tr in circuit.transformations:
                self._add(tr.gate.clone(), reg.get_shift() + tr.target,
                          tr.controls.new_with_control(control, reg.get_shift()))

    # -- execution -----------------------------------------------------------
    def _ops(self):
        """Encode the transformations as the C ABI's spz_op records."""
        n = len(self.transformations)
        arr = (N.Op_t * max(n, 1))()
        keep = []
        for i, tr in enumerate(self.transformations):
            o = arr[i]
            g = tr.gate
            o.kind = g.kind
            o.target = tr.target
            o.theta, o.phi, o.lam = g.theta, g.phi, g.lam
            o.u = 1.0
            c = tr.controls
            if g.kind == N.SWAP:
                o.target, o.t1 = g.t0, g.t1
            if c.kind == Controls.SINGLE:
                o.ones = 1 << c.controls[0]
                o.flags = N.OPF_CLASSICAL
            elif c.kind == Controls.ONES:
                o.ones = _mask(c.controls)
            elif c.kind == Controls.MIXED:
                # mc_apply drops the zero set from the mask (gates.rs:300-306)
                o.ones = _mask(q for q in c.controls if q not in c.zeros)
            if g.kind in (N.M, N.BITFLIP):
                o.u = self.rng.uniform()
            if g.kind == N.UNITARY:
                u = g.unitary
                if u.height != u.width:
                    raise ValueError("unitary must be square")
                re = np.ascontiguousarray(u.reals, dtype=np.float64)
                im = np.ascontiguousarray(u.imags, dtype=np.float64)
                keep += [re, im]
                o.m = int(u.height).bit_length() - 1
                o.u_re = re.ctypes.data
                o.u_im = im.ctypes.data
        return arr, n, keep

    def to_op_dicts(self):
        """The op list as plain dicts (kind, target, theta, phi, lam, t1,
        ones, zeros, classical, matrix) without uniform draws."""
        out = []
        for tr in self.transformations:
            g, c = tr.gate, tr.controls
            d = {"kind": g.kind, "target": tr.target, "theta": g.theta, "phi": g.phi,
                 "lam": g.lam, "t1": g.t1, "ones": 0, "zeros": 0, "classical": 0}
            if g.kind == N.SWAP:
                d["target"] = g.t0
            if c.kind == Controls.SINGLE:
                d["ones"] = 1 << c.controls[0]
                d["classical"] = 1
            elif c.kind == Controls.ONES:
                d["ones"] = _mask(c.controls)
            elif c.kind == Controls.MIXED:
                d["ones"] = _mask(q for q in c.controls if q not in c.zeros)
            if g.kind == N.UNITARY:
                d["matrix"] = g.unitary.to_matrix()
            out.append(d)
        return out

    def execute(self, mode=None):
        """QuantumCircuit::execute (circuit.rs:552-600): run and drain the
        op list.  ``mode`` selects the executor (EXEC_AUTO fused sweeps,
        EXEC_PLAIN one launch per gate, EXEC_STRICT fused in program order)."""
        arr, n, keep = self._ops()
        opts = N.ExecOpts_t()
        opts.mode = self.exec_mode if mode is None else mode
        outcomes = (ctypes.c_uint8 * max(n, 1))()
        try:
            N.check(N.lib().spz_execute(self.state._h, arr, n,
                                        ctypes.byref(self.qubit_tracker.t),
                                        ctypes.byref(opts), outcomes))
        finally:
            del keep
        self.last_outcomes = [int(outcomes[i]) for i in range(n)]
        self.transformations = []
