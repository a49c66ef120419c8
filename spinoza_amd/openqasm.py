"""OpenQASM 2.0 loader: ``load(path)`` / ``loads(text)`` -> QuantumCircuit
(openqasm.rs:13-33).

The text is parsed by the library's loader (``spz_qasm_parse``, host code in
``csrc/spz_qasm.cpp``) into the C ABI's op records, which become the circuit's
transformations exactly as the reference's builders would add them
(openqasm.rs:36-171).  ``ops_from_qasm`` returns the raw op dicts, the wire
format the executor and the oracle take.
"""
import ctypes

from . import _native as N
from .circuit import Controls, QuantumCircuit, QuantumRegister
from .gates import Gate


def _parse(text):
    if isinstance(text, str):
        text = text.encode()
    L = N.load()
    n_ops, nq = ctypes.c_uint64(), ctypes.c_int()
    N.check(L.spz_qasm_parse(text, len(text), None, 0, ctypes.byref(n_ops), ctypes.byref(nq)))
    arr = (N.Op_t * max(n_ops.value, 1))()
    N.check(L.spz_qasm_parse(text, len(text), arr, n_ops.value, ctypes.byref(n_ops),
                             ctypes.byref(nq)))
    return nq.value, [arr[i] for i in range(n_ops.value)]


def ops_from_qasm(text):
    """(qubits, [op dict]) of an OpenQASM 2.0 program."""
    nq, ops = _parse(text)
    out = [{"kind": o.kind, "target": o.target, "theta": o.theta, "phi": o.phi, "lam": o.lam,
            "t1": o.t1, "ones": o.ones, "zeros": o.zeros,
            "classical": bool(o.flags & N.OPF_CLASSICAL)} for o in ops]
    return nq, out


def loads(text, precision="f64"):
    """openqasm::loads (openqasm.rs:26-33)."""
    nq, ops = _parse(text)
    qc = QuantumCircuit(QuantumRegister(nq), precision=precision)
    for o in ops:
        g = Gate(o.kind, o.theta, o.phi, o.lam)
        ctrl = Controls.single((o.ones & -o.ones).bit_length() - 1) if o.ones else None
        qc._add(g, o.target, ctrl)
    return qc


def load(path, precision="f64"):
    """openqasm::load (openqasm.rs:13-23)."""
    with open(path, "rb") as f:
        return loads(f.read(), precision)
