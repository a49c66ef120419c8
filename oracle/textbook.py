"""Independent dense numpy simulator, written from the gate definitions.

Test infrastructure only.  It shares no code or loop structure with the C
oracle: gates are 2x2 complex matrices taken from their textbook definitions
(and the conventions of Gate::to_matrix, spinoza/src/gates.rs:95-190), applied
by tensor contraction, and controls select the rows whose control bits match.
It pins the C oracle on small states, and ``full_unitary`` gives a brute-force
2^n x 2^n matrix for tiny n (the role of ``apply_unitary`` in
spinoza/src/unitaries.rs:131-162).
"""
import numpy as np

SQ = 1.0 / np.sqrt(2.0)


def gate_matrix(kind, theta=0.0, phi=0.0, lam=0.0):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    if kind == "H":
        return np.array([[SQ, SQ], [SQ, -SQ]], dtype=complex)
    if kind == "X":
        return np.array([[0, 1], [1, 0]], dtype=complex)
    if kind == "Y":
        return np.array([[0, -1j], [1j, 0]], dtype=complex)
    if kind == "Z":
        return np.array([[1, 0], [0, -1]], dtype=complex)
    if kind == "P":
        return np.array([[1, 0], [0, np.exp(1j * theta)]], dtype=complex)
    if kind == "RX":
        return np.array([[c, -1j * s], [-1j * s, c]], dtype=complex)
    if kind == "RY":
        return np.array([[c, -s], [s, c]], dtype=complex)
    if kind == "RZ":
        return np.array([[np.exp(-0.5j * theta), 0], [0, np.exp(0.5j * theta)]], dtype=complex)
    if kind == "U":
        return np.array([[c, -np.exp(1j * lam) * s],
                         [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]], dtype=complex)
    raise ValueError(kind)


def basis(n, index=0):
    psi = np.zeros(1 << n, dtype=complex)
    psi[index] = 1.0
    return psi


def _index_bits(n):
    return np.arange(1 << n, dtype=np.int64)


def apply(psi, n, G, target, ones=(), zeros=()):
    """Return (controlled) G applied to `target` of the length-2^n vector psi."""
    psi = np.asarray(psi, dtype=complex)
    v = psi.reshape(1 << (n - 1 - target), 2, 1 << target)
    new = np.einsum("ij,ajb->aib", G, v).reshape(-1)
    if not ones and not zeros:
        return new
    idx = _index_bits(n)
    sel = np.ones(1 << n, dtype=bool)
    for q in ones:
        sel &= ((idx >> q) & 1) == 1
    for q in zeros:
        sel &= ((idx >> q) & 1) == 0
    return np.where(sel, new, psi)


def swap(psi, n, a, b):
    idx = _index_bits(n)
    ba = (idx >> a) & 1
    bb = (idx >> b) & 1
    j = idx ^ ((ba ^ bb) << a) ^ ((ba ^ bb) << b)
    return np.asarray(psi)[j]


def apply_block(psi, n, U, t, control=None):
    """Apply a 2^m x 2^m matrix to qubits [t, t+m) (qubit t = least
    significant bit of the block index)."""
    m = int(round(np.log2(U.shape[0])))
    v = np.asarray(psi, dtype=complex).reshape(1 << (n - m - t), 1 << m, 1 << t)
    new = np.einsum("ij,ajb->aib", U, v).reshape(-1)
    if control is None:
        return new
    idx = _index_bits(n)
    return np.where(((idx >> control) & 1) == 1, new, psi)


def full_unitary(n, ops):
    """Dense 2^n x 2^n matrix of a list of (kind, target, theta, phi, lam,
    ones, zeros) by applying the ops to every basis column."""
    N = 1 << n
    cols = []
    for k in range(N):
        psi = basis(n, k)
        for (kind, t, th, ph, la, ones, zeros) in ops:
            psi = apply(psi, n, gate_matrix(kind, th, ph, la), t, ones, zeros)
        cols.append(psi)
    return np.stack(cols, axis=1)


def ising_matrix_apply(psi, n, hx=1.0):
    """H psi for H = sum_{i<n-1} Z_i Z_{i+1} + hx sum_i X_i (open chain)."""
    idx = _index_bits(n)
    out = np.zeros_like(psi, dtype=complex)
    for i in range(n - 1):
        s = 1 - 2 * (((idx >> i) ^ (idx >> (i + 1))) & 1)
        out += s * psi
    for i in range(n):
        out += hx * psi[idx ^ (1 << i)]
    return out


def qft_matrix_dft(psi):
    """QFT|x> = 1/sqrt(N) sum_k e^{+2 pi i x k / N}|k>, i.e. sqrt(N) * ifft."""
    N = len(psi)
    return np.fft.ifft(psi) * np.sqrt(N)
