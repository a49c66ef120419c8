"""Dense unitaries (spinoza/src/unitaries.rs:10-162).

A ``Unitary`` is a row-major complex matrix held on the host; blocks of up to
five qubits are applied on the GPU by ``gates.transform_u``.  ``apply_unitary``
(a full 2^n x 2^n matvec) is kept for API parity as a host routine: it is the
brute-force check the reference's tests use, not a GPU workload.
"""
import numpy as np


class Unitary:
    def __init__(self, reals, imags, height, width):
        self.reals = np.ascontiguousarray(reals, dtype=np.float64).reshape(-1)
        self.imags = np.ascontiguousarray(imags, dtype=np.float64).reshape(-1)
        self.height = int(height)
        self.width = int(width)
        if len(self.reals) != self.height * self.width or len(self.imags) != len(self.reals):
            raise ValueError("matrix data does not match height x width")

    @classmethod
    def from_matrix(cls, m):
        m = np.asarray(m, dtype=complex)
        return cls(m.real.copy(), m.imag.copy(), m.shape[0], m.shape[1])

    def to_matrix(self):
        return (self.reals + 1j * self.imags).reshape(self.height, self.width)

    def clone(self):
        return Unitary(self.reals.copy(), self.imags.copy(), self.height, self.width)

    @classmethod
    def from_single_qubit_gate(cls, state, gate, target):
        """unitaries.rs:37-68: the full-state matrix of one gate."""
        n = state.n if hasattr(state, "n") else int(state)
        g = gate.to_matrix()
        dim = 1 << n
        m = np.zeros((dim, dim), dtype=complex)
        d = 1 << target
        for i in range(dim >> 1):
            s0 = i + ((i >> target) << target)
            s1 = s0 + d
            m[s0, s0], m[s0, s1], m[s1, s0], m[s1, s1] = g[0, 0], g[0, 1], g[1, 0], g[1, 1]
        return cls.from_matrix(m)

    def conj_t(self):
        """In-place conjugate transpose (unitaries.rs:71-82)."""
        m = self.to_matrix().conj().T
        self.reals = np.ascontiguousarray(m.real).reshape(-1)
        self.imags = np.ascontiguousarray(m.imag).reshape(-1)
        self.height, self.width = self.width, self.height

    def multiply(self, other):
        """unitaries.rs:85-128."""
        if self.width != other.height:
            raise ValueError("shape mismatch")
        return Unitary.from_matrix(self.to_matrix() @ other.to_matrix())

    def __str__(self):
        m = self.to_matrix()
        return "\n".join(" ".join(f"{z.real}+i{z.imag}" for z in row) for row in m) + "\n"


def apply_unitary(state, unitary):
    """unitaries.rs:131-162: full-state matvec, returns a new host (re, im)."""
    re, im = state.download()
    v = (re + 1j * im)
    out = unitary.to_matrix() @ v
    return out.real.copy(), out.imag.copy()
