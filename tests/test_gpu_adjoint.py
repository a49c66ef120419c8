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


@pytest.mark.timeout(600)
def test_batched_adjoint_config_vs_oracle(spz, orc):
    """The batched configuration's circuits (RY/RZ + CX-ladder ansatz, n in
    10..20, depth in 2..12, Ising H with hx = 1) against the oracle: every
    energy within 1e-12 and two seeded gradient entries per circuit within
    1e-10 of oracle parameter shift.  wg_max_qubits = 18 puts 10..18 qubits in
    the persistent kernel (10..12 in one workgroup each, 13..18 on groups of
    2..64 workgroups of one XCD: with the default two resident workgroups per
    CU a 64-member group fills its XCD) and 19..20 on the single-circuit path;
    the call runs with and without the graph, with as many resident workgroups
    per CU as fit, and once more with every circuit of the persistent kernel in
    one workgroup."""
    print("case batched adjoint vs oracle", flush=True)
    from spinoza_amd.autograd import batch_adjoint
    from spinoza_amd.rng import SplitMix64
    from spinoza_amd.workloads import hea
    orc.set_threads(16)
    rng = SplitMix64(5050)
    circuits, wheres = [], []
    for n in range(10, 21):
        for rep in range(2):
            depth = 2 + rng.next_u64() % 11
            ops, _, where = hea(n, depth, seed=7000 + 10 * n + rep)
            circuits.append((n, ops))
            wheres.append(where)
    results = [batch_adjoint(circuits, 1.0, wg_max_qubits=18, use_graph=g) for g in (True, False)]
    results.append(batch_adjoint(circuits, 1.0, wg_max_qubits=18, wg_per_cu=-1))
    for energies, grads, stats in results:
        assert stats["n_small"] == 18 and stats["n_large"] == 4
        assert stats["n_group"] == 12, stats
    results.append(batch_adjoint(circuits, 1.0, wg_max_qubits=18, group_qubits=-1, wg_per_cu=-1))
    assert results[-1][2]["n_group"] == 0
    for c, ((n, ops), where) in enumerate(zip(circuits, wheres)):
        picks = sorted({where[rng.next_u64() % len(where)] for _ in range(2)})
        # oracle: energy, then the shifted energies from a shared prefix
        st = orc.OracleState(n)
        done = 0
        ref_g = {}
        for i in picks:
            orc.execute(st, ops[done:i])
            done = i
            ev = []
            for sgn in (1.0, -1.0):
                s2 = st.copy()
                o = dict(ops[i])
                o["theta"] += sgn * math.pi / 2
                orc.execute(s2, [o] + ops[i + 1:])
                ev.append(s2.ising_energy(1.0))
            ref_g[i] = 0.5 * (ev[0] - ev[1])
        orc.execute(st, ops[done:])
        e_ref = st.ising_energy(1.0)
        for energies, grads, _ in results:
            assert abs(energies[c] - e_ref) <= 1e-12, (n, energies[c], e_ref)
            for i, v in ref_g.items():
                assert abs(grads[c][i] - v) <= 1e-10, (n, i, grads[c][i], v)
            others = [k for k in range(len(ops)) if k not in set(where)]
            assert np.all(grads[c][others] == 0.0)


def test_batched_adjoint_rejects_like_single(spz):
    """One validation rule for both paths: a control on the target, a qubit
    both a 1- and a 0-control, a SWAP and eight controls are rejected whatever
    the switch point puts the circuit on."""
    print("case batched adjoint validation", flush=True)
    from spinoza_amd import SpinozaError
    from spinoza_amd import _native as N
    from spinoza_amd.autograd import adjoint_gradient, batch_adjoint
    from spinoza_amd.workloads import hea, op
    good, _, _ = hea(12, 2, seed=3)
    bad_ops = [op(N.RY, 3, theta=0.4, ones=1 << 3),
               op(N.RY, 3, theta=0.4, ones=1 << 5, zeros=1 << 5),
               op(N.SWAP, 2, t1=4),
               op(N.RZ, 11, theta=0.1, ones=0xFF)]
    for bad in bad_ops:
        for wg_max in (20, 4):  # persistent kernel / single-circuit path
            with pytest.raises(SpinozaError):
                batch_adjoint([(12, good), (12, good[:5] + [bad] + good[5:])], 1.0,
                              wg_max_qubits=wg_max)
    # the library still works after the rejections
    energies, grads, _ = batch_adjoint([(12, good)], 1.0, wg_max_qubits=20)
    e1, g1, _ = adjoint_gradient(spz.State(12), good, 1.0)
    assert abs(energies[0] - e1) < 1e-12 and np.max(np.abs(grads[0] - g1)) < 1e-12


def test_adjoint_many_checkpoints_and_error_return(spz, orc):
    """Regression test for the checkpoint store: with a checkpoint after
    every op the store grows far past its first capacity, and every restore
    reads its own state (a store of handles that copied shallowly on growth
    once freed a checkpoint that a later restore then read).  Then an error
    return (a bad op past the first ones) leaves the library usable."""
    print("case adjoint many checkpoints", flush=True)
    from spinoza_amd import SpinozaError
    from spinoza_amd import _native as N
    from spinoza_amd.autograd import adjoint_gradient
    from spinoza_amd.workloads import hea, op
    n = 13
    ops, _, where = hea(n, 3, seed=31)
    e0, g0, _ = adjoint_gradient(spz.State(n), ops, 1.0, 0)
    for ck in (1, 2, 5):
        e1, g1, drift = adjoint_gradient(spz.State(n), ops, 1.0, ck)
        assert drift < 1e-12
        assert abs(e1 - e0) < 1e-13 and np.max(np.abs(g1 - g0)) < 1e-12
    assert abs(e0 - oracle_energy(orc, n, ops)) < 1e-12
    bad = ops[:40] + [op(N.M, 2)] + ops[40:]
    with pytest.raises(SpinozaError):
        adjoint_gradient(spz.State(n), bad, 1.0, 3)
    bad = ops[:40] + [op(N.RX, n + 3, theta=0.1)] + ops[40:]
    with pytest.raises(SpinozaError):
        adjoint_gradient(spz.State(n), bad, 1.0, 3)
    e2, g2, _ = adjoint_gradient(spz.State(n), ops, 1.0, 4)
    assert abs(e2 - e0) < 1e-13 and np.max(np.abs(g2 - g0)) < 1e-12


def test_states_alive_at_interpreter_exit(tmp_path):
    """States still referenced when the interpreter exits (a module global, a
    stored traceback) are left to the process's teardown: the exit is clean."""
    import os
    import subprocess
    import sys
    print("case exit with live states", flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    prog = tmp_path / "live.py"
    prog.write_text(
        "import sys\n"
        f"sys.path.insert(0, {root!r})\n"
        "import spinoza_amd as spz\n"
        "from spinoza_amd.autograd import adjoint_gradient\n"
        "from spinoza_amd.workloads import hea\n"
        "KEEP = [spz.State(12), spz.State(10, 'f32')]\n"
        "ops, _, _ = hea(12, 2, seed=1)\n"
        "adjoint_gradient(KEEP[0], ops, 1.0, 5)\n"
        "try:\n"
        "    spz.State(12).copy_from(spz.State(11))\n"
        "except spz.SpinozaError:\n"
        "    TB = sys.exc_info()\n"
        "print('done', flush=True)\n")
    r = subprocess.run([sys.executable, str(prog)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "done" in r.stdout


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
