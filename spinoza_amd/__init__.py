"""spinoza_amd -- a GPU-resident (MI355X / gfx950) engine for the spinoza
state-vector API.

Submodules mirror the reference crate: ``gates``, ``core``, ``measurement``,
``circuit``, ``unitaries``, ``math``; ``spynoza`` mirrors the Python bindings.
All compute goes through the C ABI of ``libspinoza_amd.so`` (hand-written HIP
for gfx950); there is no CPU fallback.
"""
from . import _native
from ._native import SpinozaError

_native.load()

from .core import (State, Reservoir, expect_zz, inner, iqft, ising_energy,  # noqa: E402
                   qubit_expectation_value, reservoir_sampling, sample,
                   xyz_expectation_value)
from .gates import (Gate, apply, c_apply, c_transform_u, cc_apply, mc_apply,  # noqa: E402
                    mc_apply_masked, transform_u)
from .measurement import measure_qubit  # noqa: E402
from .circuit import (Controls, QuantumCircuit, QuantumRegister,  # noqa: E402
                      QuantumTransformation)
from .unitaries import Unitary, apply_unitary  # noqa: E402

__all__ = [
    "State", "Gate", "apply", "c_apply", "cc_apply", "mc_apply", "mc_apply_masked",
    "transform_u", "c_transform_u", "measure_qubit", "qubit_expectation_value",
    "xyz_expectation_value", "expect_zz", "ising_energy", "inner", "iqft", "sample",
    "Reservoir", "reservoir_sampling", "QuantumCircuit", "QuantumRegister", "Controls",
    "QuantumTransformation", "Unitary", "apply_unitary", "SpinozaError",
]
