"""Adjoint gradients on the GPU against gradients computed by the CPU oracle.

Parameter shift (E(theta + pi/2) - E(theta - pi/2)) / 2 is exact for RX, RY,
RZ and P, so it pins the adjoint gradient to rounding (bar: 1e-10 max abs).
Controlled rotations have no two-term shift rule; they are checked against
central finite differences of the oracle energy (h = 1e-5, bar 1e-8).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_energy(orc, n, ops, hx=1.0):
    s = orc.OracleState(n)
    orc.execute(s, ops)
    return s.ising_energy(hx)


def oracle_param_shift(orc, n, ops, param_ops, hx=1.0):
    g = []
    for i in param_ops:
        plus = [dict(o) for o in ops]
        minus = [dict(o) for o in ops]
        plus[i]["theta"] += math.pi / 2
        minus[i]["theta"] -= math.pi / 2
        g.append(0.5 * (oracle_energy(orc, n, plus, hx) - oracle_energy(orc, n, minus, hx)))
    return np.array(g)


@pytest.fixture(scope="module")
def spz():
    import spinoza_amd
    return spinoza_amd


@pytest.mark.parametrize("n,layers", [(6, 2), (11, 3)])
def test_adjoint_vs_oracle_param_shift(spz, orc, n, layers):
    print(f"case adjoint hea n={n} layers={layers}", flush=True)
    from spinoza_amd.autograd import adjoint_gradient
    from spinoza_amd.workloads import hea
    ops, thetas, where = hea(n, layers, seed=40 + n)
    s = spz.State(n)
    e, g, _ = adjoint_gradient(s, ops, 1.0)
    assert abs(e - oracle_energy(orc, n, ops)) < 1e-12
    ref = oracle_param_shift(orc, n, ops, where)
    got = np.array([g[i] for i in where])
    assert np.max(np.abs(got - ref)) < 1e-10, np.max(np.abs(got - ref))
    # non-parameter ops get zero
    others = [i for i in range(len(ops)) if i not in set(where)]
    assert np.all(g[others] == 0.0)


def test_adjoint_16q_vs_oracle(spz, orc):
    """The full 96-parameter gradient at 16 qubits against oracle parameter shift."""
    print("case adjoint 16q", flush=True)
    from spinoza_amd.autograd import adjoint_gradient
    from spinoza_amd.workloads import hea
    orc.set_threads(16)
    n, layers = 16, 3
    ops, thetas, where = hea(n, layers, seed=4)
    s = spz.State(n)
    e, g, _ = adjoint_gradient(s, ops, 1.0, ckpt_every=len(ops) // 3)
    assert abs(e - oracle_energy(orc, n, ops)) < 1e-12
    ref = oracle_param_shift(orc, n, ops, where)
    got = np.array([g[i] for i in where])
    assert np.max(np.abs(got - ref)) < 1e-10, np.max(np.abs(got - ref))


def test_adjoint_mixed_gates_vs_finite_differences(spz, orc):
    print("case adjoint mixed", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.autograd import adjoint_gradient
    from spinoza_amd.workloads import op
    n = 7
    rng = np.random.default_rng(3)
    ops = []
    for layer in range(3):
        for q in range(n):
            ops.append(op(N.H, q))
            ops.append(op([N.RX, N.RY, N.RZ, N.P][(q + layer) % 4], q, theta=rng.uniform(0, 6.28)))
        ops.append(op(N.RX, 3, theta=0.4, ones=1 << 1))            # CRX
        ops.append(op(N.RY, 0, theta=1.3, ones=1 << 5))            # CRY
        ops.append(op(N.RZ, 6, theta=-0.7, ones=(1 << 2) | (1 << 4)))  # CCRZ
        ops.append(op(N.P, 2, theta=0.9, ones=1 << 6))             # CP
        ops.append(op(N.RY, 4, theta=0.2, zeros=1 << 0))           # 0-controlled RY
        ops.append(op(N.U, 5, theta=0.3, phi=0.8, lam=-0.5))
        ops.append(op(N.SWAP, 1, t1=5))
        ops.append(op(N.X, 3, ones=1 << 2))
        ops.append(op(N.Y, 1))
        ops.append(op(N.Z, 4))
    s = spz.State(n)
    e, g, drift = adjoint_gradient(s, ops, 0.7, ckpt_every=9)
    assert drift < 1e-13
    assert abs(e - oracle_energy(orc, n, ops, 0.7)) < 1e-12
    h = 1e-5
    for i, o in enumerate(ops):
        if o["kind"] not in (N.RX, N.RY, N.RZ, N.P):
            assert g[i] == 0.0
            continue
        plus = [dict(x) for x in ops]
        minus = [dict(x) for x in ops]
        plus[i]["theta"] += h
        minus[i]["theta"] -= h
        fd = (oracle_energy(orc, n, plus, 0.7) - oracle_energy(orc, n, minus, 0.7)) / (2 * h)
        assert abs(g[i] - fd) < 1e-8, (i, o, g[i], fd)


def test_adjoint_checkpoints_do_not_change_gradient(spz):
    print("case adjoint checkpoints", flush=True)
    from spinoza_amd.autograd import adjoint_gradient
    from spinoza_amd.workloads import hea
    n = 14
    ops, _, where = hea(n, 6, seed=9)
    a = spz.State(n)
    e0, g0, _ = adjoint_gradient(a, ops, 1.0, 0)
    b = spz.State(n)
    e1, g1, drift = adjoint_gradient(b, ops, 1.0, 40)
    # the checkpointed forward pass is fused in segments: rounding may differ
    assert abs(e0 - e1) < 1e-14
    assert drift < 1e-12
    assert np.max(np.abs(g0 - g1)) < 1e-12
    # the state is left holding the forward result
    assert abs(spz.ising_energy(b, 1.0) - e1) < 1e-12


def test_gpu_param_shift_matches_adjoint(spz):
    print("case gpu param shift", flush=True)
    from spinoza_amd.autograd import adjoint_gradient, param_shift_gradient
    from spinoza_amd.workloads import hea
    n = 13
    ops, _, where = hea(n, 4, seed=2)
    s = spz.State(n)
    _, g, _ = adjoint_gradient(s, ops, 1.0)
    ps = param_shift_gradient(n, ops, where, 1.0, subset=range(0, len(where), 7))
    for k, v in ps.items():
        assert abs(g[where[k]] - v) < 1e-10


def test_torch_autograd_function(spz):
    print("case torch autograd", flush=True)
    import torch
    from spinoza_amd.autograd import IsingEnergy, adjoint_gradient
    from spinoza_amd.workloads import hea
    n, layers = 9, 3

    def builder(vals):
        ops, _, where = hea(n, layers, thetas=vals)
        return ops, where

    th0 = hea(n, layers, seed=5)[1]
    th = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    e = IsingEnergy.apply(th, builder, n, 1.0)
    e.backward()
    ops, where = builder(th0)
    e2, g, _ = adjoint_gradient(spz.State(n), ops, 1.0)
    assert abs(e.item() - e2) < 1e-13
    assert np.max(np.abs(th.grad.numpy() - np.array([g[i] for i in where]))) < 1e-13


def test_adjoint_fp32_storage(spz):
    print("case adjoint fp32", flush=True)
    from spinoza_amd.autograd import adjoint_gradient
    from spinoza_amd.workloads import hea
    n = 12
    ops, _, where = hea(n, 3, seed=8)
    e64, g64, _ = adjoint_gradient(spz.State(n), ops, 1.0)
    e32, g32, _ = adjoint_gradient(spz.State(n, "f32"), ops, 1.0)
    assert abs(e64 - e32) < 1e-4
    assert np.max(np.abs(g64 - g32)) < 1e-4


def test_batched_adjoint_matches_single(spz):
    print("case batched adjoint", flush=True)
    from spinoza_amd.autograd import adjoint_gradient, batch_adjoint
    from spinoza_amd.rng import SplitMix64
    from spinoza_amd.workloads import hea
    rng = SplitMix64(55)
    circuits = []
    for i in range(40):
        n = 4 + rng.next_u64() % 9          # 4..12
        depth = 1 + rng.next_u64() % 4      # 1..4
        ops, _, _ = hea(n, depth, seed=1000 + i)
        circuits.append((n, ops))
    for wg_max, graph in ((16, True), (8, False)):
        energies, grads, stats = batch_adjoint(circuits, 0.8, wg_max_qubits=wg_max, use_graph=graph)
        assert stats["n_small"] + stats["n_large"] == len(circuits)
        for (n, ops), e, g in zip(circuits, energies, grads):
            e1, g1, _ = adjoint_gradient(spz.State(n), ops, 0.8)
            assert abs(e - e1) < 1e-12
            assert np.max(np.abs(g - g1)) < 1e-12


def test_adjoint_grouped_and_tiled_reverse_sweep(spz, orc):
    """14 qubits: long parameter-free runs (undone by the tile executor), rotation
    groups on up to 3 qubits with controls inside and outside the group, SWAPs
    and a checkpoint every 17 ops; energy against the oracle and every gradient
    against central finite differences of the oracle energy."""
    print("case adjoint grouped/tiled 14q", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.autograd import adjoint_gradient
    from spinoza_amd.workloads import op
    orc.set_threads(16)
    n = 14
    rng = np.random.default_rng(11)
    ops = []
    for layer in range(3):
        for q in range(n):
            ops.append(op(N.RY, q, theta=rng.uniform(0, 6.28)))
            ops.append(op(N.RZ, q, theta=rng.uniform(0, 6.28)))
        for q in range(0, n - 1, 2):  # parameter-free run: CX, H, CCX, U, SWAP, Y, Z
            ops.append(op(N.X, q + 1, ones=1 << q))
            ops.append(op(N.H, (q + 5) % n))
        ops.append(op(N.X, 13, ones=(1 << 2) | (1 << 9)))
        ops.append(op(N.U, 6, theta=0.3, phi=0.8, lam=-0.5))
        ops.append(op(N.SWAP, 3, t1=12))
        ops.append(op(N.Y, 7))
        ops.append(op(N.Z, 0, ones=1 << 11))
        ops.append(op(N.RX, 2, theta=0.4, ones=1 << 3))            # control inside a group
        ops.append(op(N.RY, 3, theta=1.3, ones=1 << 2))
        ops.append(op(N.RZ, 4, theta=-0.7, ones=(1 << 10) | (1 << 12)))  # controls outside
        ops.append(op(N.P, 2, theta=0.9, zeros=1 << 13))
        ops.append(op(N.RX, 9, theta=2.1, ones=1 << 0))
    s = spz.State(n)
    e, g, drift = adjoint_gradient(s, ops, 0.6, ckpt_every=17)
    assert drift < 1e-12
    assert abs(e - oracle_energy(orc, n, ops, 0.6)) < 1e-12
    h = 1e-5
    params = [i for i, o in enumerate(ops) if o["kind"] in (N.RX, N.RY, N.RZ, N.P)]
    for i in range(len(ops)):
        if i not in params:
            assert g[i] == 0.0
    for i in params[::3] + params[-6:]:
        plus = [dict(x) for x in ops]
        minus = [dict(x) for x in ops]
        plus[i]["theta"] += h
        minus[i]["theta"] -= h
        fd = (oracle_energy(orc, n, plus, 0.6) - oracle_energy(orc, n, minus, 0.6)) / (2 * h)
        assert abs(g[i] - fd) < 1e-8, (i, ops[i], g[i], fd)
