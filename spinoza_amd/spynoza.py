"""The ``spynoza`` Python surface (spynoza/src/lib.rs:21-421) over the gfx950
engine: QuantumCircuit, QuantumRegister, PyState, PyQuantumTransformation,
run, get_samples, show_table, qubit_expectation_value, xyz_expectation_value.

``QuantumCircuit.state_vector`` returns a PyState that downloads lazily and
can hand out the device arrays as torch tensors without a copy (the
reference clones the whole state, lib.rs:205-211).
"""
from . import _native as N
from . import circuit as _c
from . import core as _core

QuantumRegister = _c.QuantumRegister

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
        """(reals, imags) as torch tensors over the device memory (no copy);
        valid while the circuit's state lives and is not modified."""
        import torch
        from .torch_interop import device_views
        return device_views(self.data)


class PyQuantumTransformation:
    """lib.rs:91-170: name, target, controls, arg."""

    def __init__(self, tr):
        g = tr.gate
        if g.kind not in _GNAMES:
            raise ValueError(f"{g.name} has no Python transformation (lib.rs:159 todo!)")
        self.target = tr.target
        c = tr.controls
        if c.kind == _c.Controls.MIXED:
            raise ValueError("mixed controls have no Python form (lib.rs:147 todo!)")
        self.controls = list(c.controls) if c.kind != _c.Controls.NONE else None
        self.name = _GNAMES[g.kind]
        if g.kind in (N.P, N.RX, N.RY, N.RZ):
            self.arg = (g.theta, 0.0, 0.0)
        elif g.kind == N.U:
            self.arg = (g.theta, g.phi, g.lam)
        else:
            self.arg = None

    def __str__(self):
        s = f"name: {self.name}\ntarget: {self.target}\n"
        s += f"arg: ({self.arg[0]}, {self.arg[1]}, {self.arg[2]})\n" if self.arg else "arg: None\n"
        s += ("controls: [" + ",".join(map(str, self.controls)) + "]") if self.controls \
            else "controls: None"
        return s


class QuantumCircuit(_c.QuantumCircuit):
    """lib.rs:175-369."""

    @property
    def register_sizes(self):
        return list(self.quantum_registers_info)

    @property
    def state_vector(self):
        return PyState(self.state)

    @property
    def transformations_py(self):
        return [PyQuantumTransformation(t) for t in self.transformations]


def run(qc):
    """lib.rs:402-406: execute and return the state."""
    qc.execute()
    return qc.state_vector


def _data(s):
    return s.data if isinstance(s, PyState) else s


def xyz_expectation_value(observable, state, targets):
    return _core.xyz_expectation_value(observable, _data(state), targets)


def qubit_expectation_value(state, target):
    return _core.qubit_expectation_value(_data(state), target)


def get_samples(state, reservoir_size, num_tests=None):
    """lib.rs:393-400: histogram of reservoir_size samples."""
    return _core.reservoir_sampling(_data(state), reservoir_size, num_tests).get_outcome_count()


def show_table(state):
    """A plain-text table of index, amplitude and probability (the
    reference's colour table, utils.rs:50-88, is a display helper)."""
    re, im = _data(state).download()
    rows = [f"{i:>8}  {a:+.6f} {b:+.6f}i  {a * a + b * b:.6f}" for i, (a, b) in
            enumerate(zip(re, im))]
    return "\n".join(["   index  amplitude                    probability"] + rows)
