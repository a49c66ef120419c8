"""Seeded synthetic circuits for the benchmarks and parity tests.

Every generator returns a list of op dicts (see ``executor``) built from
SplitMix64 draws, so the same circuit can be replayed by the oracle.
"""
import math

from . import _native as N
from .rng import SplitMix64

TWO_PI = 2.0 * math.pi
SINGLE_KINDS = [N.H, N.X, N.Y, N.Z, N.RX, N.RY, N.RZ, N.P, N.U]


def op(kind, target, theta=0.0, phi=0.0, lam=0.0, t1=0, ones=0, zeros=0, classical=False):
    return {"kind": kind, "target": target, "theta": theta, "phi": phi, "lam": lam,
            "t1": t1, "ones": ones, "zeros": zeros, "classical": classical}


def qft(n):
    """Quantum Fourier transform on n qubits: H on qubit j (from the most
    significant down), controlled-phase P(pi/2^(j-k)) from every lower qubit
    k, then the qubit-reversal SWAP layer.  Maps |x> to
    N^-1/2 sum_k exp(2 pi i x k / N) |k>."""
    ops = []
    for j in range(n - 1, -1, -1):
        ops.append(op(N.H, j))
        for k in range(j - 1, -1, -1):
            ops.append(op(N.P, j, theta=math.pi / float(2 ** (j - k)), ones=1 << k))
    for i in range(n // 2):
        ops.append(op(N.SWAP, i, t1=n - 1 - i))
    return ops


def random_circuit(n, layers, seed):
    """Each layer: one gate per qubit drawn uniformly from
    {H, X, Y, Z, RX, RY, RZ, P, U} with angles uniform in [0, 2 pi), then
    floor(n/2) CX or CZ (50/50) on a seeded random perfect matching (control
    first, target second)."""
    rng = SplitMix64(seed)
    ops = []
    for _ in range(layers):
        for q in range(n):
            kind = SINGLE_KINDS[rng.next_u64() % len(SINGLE_KINDS)]
            if kind in (N.RX, N.RY, N.RZ, N.P):
                ops.append(op(kind, q, theta=rng.uniform() * TWO_PI))
            elif kind == N.U:
                ops.append(op(kind, q, theta=rng.uniform() * TWO_PI, phi=rng.uniform() * TWO_PI,
                              lam=rng.uniform() * TWO_PI))
            else:
                ops.append(op(kind, q))
        perm = list(range(n))
        for i in range(n - 1, 0, -1):  # Fisher-Yates
            j = rng.next_u64() % (i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for p in range(n // 2):
            c, t = perm[2 * p], perm[2 * p + 1]
            kind = N.X if (rng.next_u64() & 1) == 0 else N.Z
            ops.append(op(kind, t, ones=1 << c))
    return ops


def hea(n, layers, thetas=None, seed=None):
    """Hardware-efficient ansatz: per layer RY(theta) then RZ(theta) on every
    qubit, then CX(i, i+1) for i < n-1.  Returns (ops, thetas, param_ops):
    param_ops[k] is the op index of parameter k."""
    nparams = 2 * n * layers
    if thetas is None:
        rng = SplitMix64(seed if seed is not None else 4)
        thetas = [rng.uniform() * TWO_PI for _ in range(nparams)]
    ops, where = [], []
    k = 0
    for _ in range(layers):
        for q in range(n):
            where.append(len(ops))
            ops.append(op(N.RY, q, theta=thetas[k])); k += 1
            where.append(len(ops))
            ops.append(op(N.RZ, q, theta=thetas[k])); k += 1
        for q in range(n - 1):
            ops.append(op(N.X, q + 1, ones=1 << q))
    return ops, list(thetas), where


def qcbm(n):
    """The qcbm circuit of the reference's tests (spinoza/src/gates.rs:1499-1529)."""
    pairs = [(i, (i + 1) % n) for i in range(n)][: n - 1]
    ops = []
    for i in range(n):
        ops += [op(N.RX, i, theta=1.0), op(N.RZ, i, theta=1.0)]
    ops += [op(N.X, b, ones=1 << a) for a, b in pairs]
    for _ in range(9):
        for i in range(n):
            ops += [op(N.RZ, i, theta=1.0), op(N.RX, i, theta=1.0), op(N.RZ, i, theta=1.0)]
        ops += [op(N.X, b, ones=1 << a) for a, b in pairs]
    for i in range(n):
        ops += [op(N.RZ, i, theta=1.0), op(N.RX, i, theta=1.0)]
    return ops


def gate_bytes(o, n, amp_bytes=16):
    """Algorithmic HBM bytes of one op: read + write of every amplitude the
    reference's kernel changes (full 2x2: all; Z/P and single-control: half;
    CP/CZ: a quarter; RZ: all, controlled RZ: half)."""
    N_ = 1 << n
    k = o["kind"]
    nc = bin(o.get("ones", 0) | o.get("zeros", 0)).count("1")
    frac = 1.0 / (1 << nc)
    if k in (N.Z, N.P):
        frac *= 0.5
    if k == N.SWAP:
        frac *= 0.5
    return 2.0 * amp_bytes * N_ * frac
