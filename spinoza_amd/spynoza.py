"""The ``spynoza`` Python surface (spynoza/src/lib.rs:21-421) over the gfx950
engine: QuantumCircuit, QuantumRegister, PyState, PyQuantumTransformation,
QuantumTransformation, run, get_samples, show_table, qubit_expectation_value,
xyz_expectation_value.

As in the reference, ``QuantumCircuit`` wraps the crate-level circuit
(``struct QuantumCircuit { qc }``, lib.rs:173-176) and ``state_vector`` hands
out a snapshot (the reference clones the state, lib.rs:205-211): here the
snapshot is a device-to-device copy, downloaded to the host only when it is
read, and it can also be viewed as torch tensors without a host copy.
"""
from . import _native as N
from . import circuit as _c
from . import core as _core
from .gates import Gate

_GNAMES = {N.H: "h", N.X: "x", N.Y: "y", N.Z: "z", N.P: "p", N.RX: "rx", N.RY: "ry",
           N.RZ: "rz", N.U: "u"}


class PyState:
    """lib.rs:34-58: len, [i] -> (re, im), str."""

    def __init__(self, state):
        self.data = state
        self._cache = None

    def _host(self):
        if self._cache is None:
            self._cache = self.data.download()
        return self._cache

    def __len__(self):
        return len(self.data)

    def __getitem__(self, i):
        re, im = self._host()
        return float(re[i]), float(im[i])

    def __str__(self):
        re, im = self._host()
        return "".join(f"{a} + i{b}\n" for a, b in zip(re, im))

    def to_numpy(self):
        re, im = self._host()
        return re + 1j * im

    def to_torch(self):
        """(reals, imags) as torch tensors over the snapshot's device memory
        (no copy)."""
        from .torch_interop import device_views
        return device_views(self.data)


class QuantumRegister:
    """lib.rs:62-87: ``QuantumRegister(size)``, len, update_shift."""

    def __init__(self, size):
        self.qr = _c.QuantumRegister(size)

    def __len__(self):
        return len(self.qr)

    def update_shift(self, shift):
        self.qr.update_shift(shift)


class PyQuantumTransformation:
    """lib.rs:89-163: name, target, controls, arg (get and set)."""

    def __init__(self, name, target, controls=None, arg=None):
        self.name = name
        self.target = int(target)
        self.controls = list(controls) if controls is not None else None
        self.arg = tuple(arg) if arg is not None else None

    @classmethod
    def from_transformation(cls, tr):
        """From<QuantumTransformationRS> (lib.rs:136-163).  The reference
        leaves mixed controls and non-parametric special gates as todo!();
        they raise ValueError here."""
        g = tr.gate
        if g.kind not in _GNAMES:
            raise ValueError(f"{g.name} has no Python transformation (lib.rs:159 todo!)")
        c = tr.controls
        if c.kind == _c.Controls.MIXED:
            raise ValueError("mixed controls have no Python form (lib.rs:147 todo!)")
        controls = list(c.controls) if c.kind != _c.Controls.NONE else None
        if g.kind in (N.P, N.RX, N.RY, N.RZ):
            arg = (g.theta, 0.0, 0.0)
        elif g.kind == N.U:
            arg = (g.theta, g.phi, g.lam)
        else:
            arg = None
        return cls(_GNAMES[g.kind], tr.target, controls, arg)

    def __str__(self):
        s = f"name: {self.name}\ntarget: {self.target}\n"
        s += f"arg: ({self.arg[0]}, {self.arg[1]}, {self.arg[2]})\n" if self.arg else "arg: None\n"
        s += ("controls: [" + ",".join(map(str, self.controls)) + "]") if self.controls \
            else "controls: None"
        return s


class QuantumTransformation:
    """lib.rs:165-170: an opaque crate-level transformation for
    ``QuantumCircuit.add``."""

    def __init__(self, gate, target, controls=None):
        self.qt = _c.QuantumTransformation(gate, target, controls)


class QuantumCircuit:
    """lib.rs:172-369."""

    def __init__(self, *registers):
        if len(registers) == 1 and isinstance(registers[0], (list, tuple)):
            registers = tuple(registers[0])
        # the binding extracts copies of the registers (lib.rs:181-187), so
        # the caller's registers keep their shift
        copies = []
        for r in registers:
            q = _c.QuantumRegister(1)
            q.qubits = list(r.qr.qubits)
            copies.append(q)
        self.qc = _c.QuantumCircuit(*copies)

    @property
    def num_qubits(self):
        return self.qc.state.n

    @property
    def register_sizes(self):
        return list(self.qc.quantum_registers_info)

    @property
    def state_vector(self):
        return PyState(self.qc.state.clone())

    @property
    def transformations(self):
        return [PyQuantumTransformation.from_transformation(t) for t in self.qc.transformations]

    def inverse(self):
        self.qc.inverse()

    def h(self, target): self.qc.h(target)
    def x(self, target): self.qc.x(target)
    def y(self, target): self.qc.y(target)
    def z(self, target): self.qc.z(target)
    def p(self, angle, target): self.qc.p(angle, target)
    def rx(self, angle, target): self.qc.rx(angle, target)
    def ry(self, angle, target): self.qc.ry(angle, target)
    def rz(self, angle, target): self.qc.rz(angle, target)
    def u(self, theta, phi, lam, target): self.qc.u(theta, phi, lam, target)
    def ch(self, control, target): self.qc.ch(control, target)
    def cx(self, control, target): self.qc.cx(control, target)
    def ccx(self, control0, control1, target): self.qc.ccx(control0, control1, target)
    def cy(self, control, target): self.qc.cy(control, target)
    def cu(self, theta, phi, lam, control, target): self.qc.cu(theta, phi, lam, control, target)
    def cp(self, angle, control, target): self.qc.cp(angle, control, target)
    def crx(self, angle, control, target): self.qc.crx(angle, control, target)
    def cry(self, angle, control, target): self.qc.cry(angle, control, target)
    def crz(self, angle, control, target): self.qc.crz(angle, control, target)
    def measure(self, target): self.qc.measure(target)
    def swap(self, t0, t1): self.qc.swap(t0, t1)
    def iqft(self, targets): self.qc.iqft(list(targets))
    def bit_flip_noise(self, prob, target): self.qc.bit_flip_noise(prob, target)

    def add(self, q_transformation):
        t = q_transformation.qt
        self.qc.add(_c.QuantumTransformation(t.gate.clone(), t.target, t.controls.clone()))

    def append(self, circuit, quantum_register):
        self.qc.append(circuit.qc, quantum_register.qr)

    def c_append(self, circuit, c, quantum_register):
        self.qc.c_append(circuit.qc, c, quantum_register.qr)

    def mc_append(self, circuit, cs, quantum_register):
        self.qc.mc_append(circuit.qc, list(cs), quantum_register.qr)

    def execute(self):
        self.qc.execute()


def run(qc):
    """lib.rs:402-406: execute and return the state."""
    qc.execute()
    return qc.state_vector


def _data(s):
    return s.data if isinstance(s, PyState) else s


def xyz_expectation_value(observable, state, targets):
    """lib.rs:381-384."""
    return _core.xyz_expectation_value(observable, _data(state), list(targets))


def qubit_expectation_value(state, target):
    """lib.rs:386-389."""
    return _core.qubit_expectation_value(_data(state), target)


def get_samples(state, reservoir_size, num_tests):
    """lib.rs:391-400: histogram {index: count} of reservoir_size samples
    drawn from |psi|^2 (inverse-CDF sampler; num_tests is accepted for the
    reference's signature)."""
    return _core.reservoir_sampling(_data(state), reservoir_size, num_tests).get_outcome_count()


def show_table(state):
    """lib.rs:60-63 (utils::to_table): one row per basis state with its
    amplitude and probability, as plain text (the reference colours it)."""
    re, im = _data(state).download()
    rows = [f"{i:>8}  {a:+.6f} {b:+.6f}i  {a * a + b * b:.6f}" for i, (a, b) in
            enumerate(zip(re, im))]
    return "\n".join(["   index  amplitude                    probability"] + rows)


__all__ = ["PyState", "QuantumRegister", "PyQuantumTransformation", "QuantumTransformation",
           "QuantumCircuit", "run", "xyz_expectation_value", "qubit_expectation_value",
           "get_samples", "show_table", "Gate"]
