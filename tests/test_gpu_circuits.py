"""GPU parity of the circuit layer against the oracle.

* ``spz_execute`` over op lists that hold every arm of
  QuantumCircuit::execute (circuit.rs:552-600): measurements with drawn
  uniforms and the qubit tracker, a repeated measurement (a no-op), gates
  classically controlled by a measured qubit, single controls on unmeasured
  qubits, BitFlipNoise on both sides of its draw (gates.rs:1365-1374), dense
  blocks with and without a control, SWAPs and ordinary gates; in PLAIN and
  STRICT (bit-identical to the oracle) and AUTO (<= 1e-12, same outcomes).
* The reference's circuit-level tests restated on the GPU: ch / crz
  (circuit.rs:677-719), all_gates_as_transformations (:770-822), measure
  (:824-889), swap_all_qubits (:891-925), inverse_iqft (:927-960), inverse
  (:962-990), append / c_append / mc_append value encoding (:1045-1210),
  bit_flip_noise (:1212-1248), controlled_u (:1250-1266), and the dense
  blocks against the functional gates (unitaries.rs:197-340).
* The spynoza surface (spynoza/src/lib.rs:21-421): run, state_vector,
  get_samples, expectation values, transformations.
* The OpenQASM loader on the reference's fixtures (openqasm.rs:194-304).
"""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-12


@pytest.fixture(scope="module")
def spz():
    import spinoza_amd
    return spinoza_amd


def down(s):
    return s.download()


def max_err(s, ref):
    re, im = s.download()
    return max(np.max(np.abs(re - ref.re)), np.max(np.abs(im - ref.im)))


def exact(s, ref):
    re, im = s.download()
    return np.array_equal(re, ref.re) and np.array_equal(im, ref.im)


def mixed_ops(n, seed):
    """A circuit with every execute arm (see the module docstring)."""
    from spinoza_amd import _native as N
    from spinoza_amd.rng import SplitMix64
    from spinoza_amd.workloads import op, random_circuit
    rng = SplitMix64(seed)
    ops = random_circuit(n, 2, seed)
    m0, m1 = 1, n - 2
    ops += [dict(op(N.M, m0), u=rng.uniform()), dict(op(N.M, m1), u=rng.uniform()),
            dict(op(N.M, m0), u=rng.uniform())]                       # second M: no-op
    ops += random_circuit(n, 1, seed + 1)
    ops += [op(N.X, 3, ones=1 << m0, classical=True),                  # classical control
            op(N.RY, 0, theta=0.7, ones=1 << m1, classical=True),
            op(N.H, 5, ones=1 << 4, classical=True),                   # 4 unmeasured: c_apply
            op(N.U, 6, 0.3, 1.1, 2.2, ones=(1 << 2) | (1 << 7), zeros=1 << 8),
            dict(op(N.BITFLIP, 2, theta=0.25), u=0.2),                 # flips
            dict(op(N.BITFLIP, 9, theta=0.25), u=0.3),                 # does not
            dict(op(N.BITFLIP, 4, theta=0.5), u=0.5),                  # eps == p flips
            op(N.SWAP, 0, t1=n - 1)]
    rs = np.random.default_rng(seed)
    for m, t, c in [(2, 3, -1), (3, n - 4, 1), (1, 0, n - 1)]:
        a = rs.normal(size=(1 << m, 1 << m)) + 1j * rs.normal(size=(1 << m, 1 << m))
        u, _ = np.linalg.qr(a)
        d = dict(op(N.UNITARY, t, ones=(1 << c) if c >= 0 else 0), matrix=u)
        ops.append(d)
    ops += random_circuit(n, 2, seed + 2)
    ops += [dict(op(N.M, 7), u=rng.uniform()), op(N.Z, 8, ones=1 << 7, classical=True)]
    ops += random_circuit(n, 1, seed + 3)
    return ops


@pytest.mark.parametrize("mode", ["plain", "strict", "auto"])
@pytest.mark.parametrize("n", [10, 13, 16])
def test_execute_every_arm_vs_oracle(spz, orc, mode, n):
    """Every arm of execute, PLAIN/STRICT bit-identical, AUTO <= 1e-12."""
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    print(f"case execute arms n={n} mode={mode}", flush=True)
    ops = mixed_ops(n, 40 + n)
    ref = orc.OracleState(n)
    trk_ref = orc.execute(ref, ops)
    s = spz.State(n)
    tr = N.Tracker_t(0, 0)
    m = {"plain": N.EXEC_PLAIN, "strict": N.EXEC_STRICT, "auto": N.EXEC_AUTO}[mode]
    outcomes = run_ops(s, ops, mode=m, tracker=tr)
    assert (tr.measured, tr.values) == (int(trk_ref[0]), int(trk_ref[1]))
    # fresh measurements report their outcome, the repeated one 255
    mpos = [i for i, o in enumerate(ops) if o["kind"] == N.M]
    assert outcomes[mpos[2]] == 255
    for i in (mpos[0], mpos[1], mpos[3]):
        q = ops[i]["target"]
        assert outcomes[i] == (int(trk_ref[1]) >> q) & 1
    err = max_err(s, ref)
    assert err <= TOL, err
    if mode != "auto":
        assert exact(s, ref), f"{mode}: not bit-identical (max err {err})"


def test_execute_fp32_strict_within_rounding(spz, orc):
    """FP32 storage: STRICT agrees with PLAIN within FP32 rounding (a sweep
    rounds to FP32 once, PLAIN after every gate), both near the FP64 oracle."""
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import random_circuit
    print("case fp32 strict", flush=True)
    n = 14
    ops = random_circuit(n, 6, 77)
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    out = {}
    for mode in (N.EXEC_PLAIN, N.EXEC_STRICT):
        s = spz.State(n, "f32")
        run_ops(s, ops, mode=mode)
        out[mode] = s.download()
    d = max(np.max(np.abs(out[N.EXEC_PLAIN][0] - out[N.EXEC_STRICT][0])),
            np.max(np.abs(out[N.EXEC_PLAIN][1] - out[N.EXEC_STRICT][1])))
    assert d < 5e-6, d
    for mode in out:
        e = max(np.max(np.abs(out[mode][0] - ref.re)), np.max(np.abs(out[mode][1] - ref.im)))
        assert e < 5e-6, (mode, e)


# ---- the reference's circuit tests, restated --------------------------------

def test_ch_crz_circuit_equals_functional(spz):
    """circuit.rs:677-719."""
    print("case ch crz", flush=True)
    n = 3
    for build, gate in [(lambda qc: qc.ch(0, 1), spz.Gate.H),
                        (lambda qc: qc.crz(math.pi / 2.0, 0, 1), spz.Gate.RZ(math.pi / 2.0))]:
        st = spz.State(n)
        for t in range(n):
            spz.apply(spz.Gate.H, st, t)
        spz.c_apply(gate, st, 0, 1)
        qc = spz.QuantumCircuit(spz.QuantumRegister(n))
        for t in range(n):
            qc.h(t)
        build(qc)
        qc.execute()
        a, b = down(qc.state), down(st)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("mode", ["plain", "strict", "auto"])
def test_all_gates_as_transformations(spz, orc, mode):
    """circuit.rs:770-822 at N = 17 (the executor tiles at this size)."""
    from spinoza_amd import _native as N
    print(f"case all gates {mode}", flush=True)
    PI = math.pi
    n = 17
    qc = spz.QuantumCircuit(spz.QuantumRegister(n))
    for t in range(n):
        qc.h(t)
    qc.x(0); qc.y(1); qc.z(2); qc.p(PI, 3); qc.cp(PI, 3, 4); qc.rx(PI, 5); qc.ry(PI, 6)
    qc.rz(PI, 7); qc.u(PI, PI, PI, 8); qc.cy(9, 10); qc.crx(PI, 11, 12); qc.cry(PI, 13, 14)
    ops = qc.to_op_dicts()
    qc.execute(mode={"plain": N.EXEC_PLAIN, "strict": N.EXEC_STRICT, "auto": N.EXEC_AUTO}[mode])
    st = spz.State(n)
    G = spz.Gate
    for t in range(n):
        spz.apply(G.H, st, t)
    spz.apply(G.X, st, 0); spz.apply(G.Y, st, 1); spz.apply(G.Z, st, 2)
    spz.apply(G.P(PI), st, 3); spz.c_apply(G.P(PI), st, 3, 4); spz.apply(G.RX(PI), st, 5)
    spz.apply(G.RY(PI), st, 6); spz.apply(G.RZ(PI), st, 7); spz.apply(G.U(PI, PI, PI), st, 8)
    spz.c_apply(G.Y, st, 9, 10); spz.c_apply(G.RX(PI), st, 11, 12)
    spz.c_apply(G.RY(PI), st, 13, 14)
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    assert exact(st, ref)
    assert max_err(qc.state, ref) <= TOL
    if mode != "auto":
        assert exact(qc.state, ref)


def test_measure_every_qubit_twice(spz, orc):
    """circuit.rs:824-889 at N = 21: measure all qubits, then again (no-op);
    the tracker keeps the first values, and outcomes and the collapsed state
    equal the oracle's for the same uniforms."""
    from spinoza_amd.rng import SplitMix64
    print("case measure 21q", flush=True)
    n = 21
    base = orc.random_state(21, n)
    assert abs(base.norm() - 1.0) < 1e-3
    s = spz.State(n)
    s.upload(base.re, base.im)
    qc = spz.QuantumCircuit.with_state(s)
    qc.rng = SplitMix64(2121)
    draws = SplitMix64(2121)
    for t in range(n):
        qc.measure(t)
    ops = qc.to_op_dicts()
    for o in ops:
        o["u"] = draws.uniform()
    qc.execute()
    vals = [qc.qubit_tracker.get_qubit_measured_val(t) for t in range(n)]
    ref = base.copy()
    trk = orc.execute(ref, ops)
    assert vals == [int(trk[1] >> t) & 1 for t in range(n)]
    assert exact(qc.state, ref)
    for t in range(n):
        qc.measure(t)
    qc.execute()
    for t in range(n):
        assert qc.qubit_tracker.is_qubit_measured(t)
        assert qc.qubit_tracker.get_qubit_measured_val(t) == vals[t]
    assert exact(qc.state, ref)  # the repeated measurements changed nothing
    # execute measures with reset (measurement.rs:84-86: outcome 1 is flipped
    # back), so every qubit ends in |0>
    re, im = qc.state.download()
    assert abs(re[0] ** 2 + im[0] ** 2 - 1.0) < 1e-12


def test_swap_all_qubits(spz, orc):
    """circuit.rs:891-925: circuit SWAPs equal three CX each, bit for bit."""
    print("case swap all", flush=True)
    n = 9
    base = orc.random_state(9, n)
    s = spz.State(n)
    s.upload(base.re, base.im)
    qc = spz.QuantumCircuit.with_state(s)
    st = spz.State(n)
    st.upload(base.re, base.im)
    for i in range(n >> 1):
        qc.swap(i, n - 1 - i)
        a, b = i, n - 1 - i
        spz.c_apply(spz.Gate.X, st, a, b)
        spz.c_apply(spz.Gate.X, st, b, a)
        spz.c_apply(spz.Gate.X, st, a, b)
    qc.execute()
    x, y = down(qc.state), down(st)
    assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


@pytest.mark.parametrize("n", [2, 12])
def test_inverse_iqft_round_trip(spz, orc, n):
    """circuit.rs:927-960: iqft then its inverse returns the input state."""
    print(f"case inverse iqft n={n}", flush=True)
    base = orc.random_state(500 + n, n)
    s = spz.State(n)
    s.upload(base.re, base.im)
    qc = spz.QuantumCircuit.with_state(s)
    targets = list(range(n))[::-1]
    qc.iqft(targets)
    ops1 = qc.to_op_dicts()
    qc.execute()
    ref = base.copy()
    orc.execute(ref, ops1)
    assert max_err(qc.state, ref) <= TOL
    qc.iqft(targets)
    qc.inverse()
    qc.execute()
    assert max_err(qc.state, base) <= 1e-12


def test_inverse(spz):
    """circuit.rs:962-990: inverse() of [H(0), P(pi/4, 1)] equals
    [P(-pi/4, 1), H(0)] bit for bit."""
    print("case inverse", flush=True)
    qc1 = spz.QuantumCircuit.with_state(spz.State(2))
    qc1.h(0)
    qc1.p(math.pi / 4.0, 1)
    qc1.inverse()
    qc1.execute()
    qc2 = spz.QuantumCircuit.with_state(spz.State(2))
    qc2.p(-(math.pi / 4.0), 1)
    qc2.h(0)
    qc2.execute()
    a, b = down(qc1.state), down(qc2.state)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _gate_circuit(spz, gate, n, target):
    qc = spz.QuantumCircuit(spz.QuantumRegister(n))
    qc.add(spz.QuantumTransformation(gate, target))
    return qc


@pytest.mark.parametrize("how", ["append", "c_append", "mc_append"])
def test_value_encoding_with_append(spz, orc, how):
    """circuit.rs:1045-1210: H, phases 2 pi v / 2^(t+1), then an IQFT built
    by append / c_append / mc_append encodes v = 4 as the basis state |4>."""
    print(f"case value encoding {how}", flush=True)
    PI = math.pi
    n, v = 3, 4.0
    qr = spz.QuantumRegister(n)
    qc = spz.QuantumCircuit(qr)
    for t in range(n):
        qc.append(_gate_circuit(spz, spz.Gate.H, n, t), qr)
    for t in range(n):
        qc.append(_gate_circuit(spz, spz.Gate.P(2.0 * PI / (2.0 ** (t + 1)) * v), n, t), qr)
    targets = list(range(n))[::-1]
    if how == "append":
        iq = spz.QuantumCircuit(spz.QuantumRegister(n))
        iq.iqft(targets)
    else:
        iqr = spz.QuantumRegister(n)
        iq = spz.QuantumCircuit(iqr)
        for j in range(n - 1, -1, -1):
            iq.append(_gate_circuit(spz, spz.Gate.H, n, targets[j]), iqr)
            for k in range(j - 1, -1, -1):
                r = spz.QuantumRegister(1)
                r.update_shift(targets[k])
                g = _gate_circuit(spz, spz.Gate.P(-PI / 2.0 ** (j - k)), 1, 0)
                if how == "mc_append":
                    iq.mc_append(g, [targets[j]], r)
                else:
                    iq.c_append(g, targets[j], r)
    qc.append(iq, qr)
    ops = qc.to_op_dicts()
    qc.execute()
    re, im = qc.state.download()
    want = np.zeros(1 << n)
    want[int(v)] = 1.0
    assert np.max(np.abs(re - want)) < 1e-4 and np.max(np.abs(im)) < 1e-4
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    assert exact(qc.state, ref)


def test_bit_flip_noise(spz, orc):
    """circuit.rs:1212-1248: probability 0 changes nothing, probability 1 is
    X (the reference's test builds the circuits but never executes them; here
    they run)."""
    print("case bit flip", flush=True)
    from spinoza_amd.rng import SplitMix64
    base = orc.random_state(1, 1)
    for prob, flips in [(0.0, False), (1.0, True)]:
        s = spz.State(1)
        s.upload(base.re, base.im)
        qc = spz.QuantumCircuit.with_state(s)
        qc.rng = SplitMix64(99)
        qc.bit_flip_noise(prob, 0)
        qc.execute()
        ref = base.copy()
        if flips:
            ref.apply("X", 0)
        assert exact(qc.state, ref)


def test_controlled_u(spz):
    """circuit.rs:1250-1266."""
    print("case controlled u", flush=True)
    qc = spz.QuantumCircuit(spz.QuantumRegister(3))
    qc.cu(1.0, 2.0, 3.0, 0, 1)
    qc.execute()
    st = spz.State(3)
    spz.c_apply(spz.Gate.U(1.0, 2.0, 3.0), st, 0, 1)
    a, b = down(qc.state), down(st)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


@pytest.mark.parametrize("kind", ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U"])
def test_dense_equals_functional(spz, orc, kind):
    """unitaries.rs:197-340: the full-state matrix of one gate applied as a
    dense block equals the functional gate (from |00>, as the reference
    tests, and from a random state against the oracle's dense product)."""
    print(f"case dense vs functional {kind}", flush=True)
    G = spz.Gate
    gate = {"H": G.H, "X": G.X, "Y": G.Y, "Z": G.Z, "P": G.P(3.0), "RX": G.RX(3.0),
            "RY": G.RY(3.0), "RZ": G.RZ(3.0), "U": G.U(1.0, 2.0, 3.0)}[kind]
    n = 2
    st = spz.State(n)
    u = spz.Unitary.from_single_qubit_gate(st, gate, 0)
    dense = spz.State(n)
    spz.transform_u(dense, u, 0)
    func = spz.State(n)
    spz.apply(gate, func, 0)
    a, b = down(dense), down(func)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    base = orc.random_state(7, n)
    dense.upload(base.re, base.im)
    spz.transform_u(dense, u, 0)
    ref = base.copy()
    ref.transform_u(u.to_matrix(), 0)
    assert exact(dense, ref)


# ---- spynoza ------------------------------------------------------------------

def test_spynoza_surface(spz, orc):
    """spynoza/src/lib.rs:21-421 over the GPU: value encoding through the
    binding's QuantumCircuit, run(), a state_vector snapshot that later
    executes do not change, expectation values, samples, transformations."""
    from spinoza_amd import spynoza as sp
    print("case spynoza", flush=True)
    n, v = 4, 2.4
    q = sp.QuantumRegister(n)
    qc = sp.QuantumCircuit(q)
    for i in range(n):
        qc.h(i)
    for i in range(n):
        qc.p(2 * math.pi / (2 ** (i + 1)) * v, i)
    qc.iqft(range(n)[::-1])
    tr = qc.transformations
    assert [t.name for t in tr[:n]] == ["h"] * n and tr[n].arg[0] == 2 * math.pi / 2 * v
    assert str(tr[-1]).startswith("name: h")
    assert "controls: [" in str(tr[n + n + 1])
    ops = qc.qc.to_op_dicts()
    state = sp.run(qc)
    assert len(state) == 1 << n and qc.num_qubits == n and qc.register_sizes == [n]
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    for i in range(1 << n):
        assert state[i] == (ref.re[i], ref.im[i])
    assert str(state).splitlines()[0] == f"{ref.re[0]} + i{ref.im[0]}"
    # a snapshot does not follow later executes (lib.rs:205-211 clones);
    # `snap` is read only after the next execute
    snap = qc.state_vector
    qc.x(0)
    qc.execute()
    assert snap[0] == (ref.re[0], ref.im[0])
    assert qc.state_vector[1] == (ref.re[0], ref.im[0])
    # expectation values through the binding's functions
    e2 = sp.qubit_expectation_value(state, 2)
    assert e2 == ref.qubit_expectation_value(2)
    assert sp.xyz_expectation_value("z", state, [0, 1, 2, 3]) == \
        ref.xyz_expectation_value("z", [0, 1, 2, 3])
    assert sp.xyz_expectation_value("x", state, [1]) == ref.xyz_expectation_value("x", [1])
    counts = sp.get_samples(state, 4096, 100)
    assert sum(counts.values()) == 4096
    p = ref.re ** 2 + ref.im ** 2
    top = int(np.argmax(p))
    assert counts.get(top, 0) > 4096 * p[top] * 0.8
    assert "probability" in sp.show_table(state).splitlines()[0]
    # the expectation example (spynoza/examples/expectation.py): <Z0> = cos(.54) cos(.12)
    qc2 = sp.QuantumCircuit(sp.QuantumRegister(3))
    qc2.rx(0.54, 0)
    qc2.ry(0.12, 0)
    qc2.execute()
    z = sp.xyz_expectation_value("z", qc2.state_vector, [0, 1, 2])
    assert abs(z[0] - math.cos(0.54) * math.cos(0.12)) < 1e-12
    assert abs(z[1] - 1.0) < 1e-15 and abs(z[2] - 1.0) < 1e-15


# ---- OpenQASM -----------------------------------------------------------------

def _qasm(name):
    return os.path.join(ROOT, "tests", "golden", "qasm", name)


@pytest.mark.parametrize("by_text", [False, True])
def test_qasm_iqft(spz, orc, by_text):
    """openqasm.rs:215-259 (iqft_load / iqft_loads): the program's state
    equals the circuit iqft of the same width (both from |0000>), and the
    loaded circuit matches the oracle on its ops bit for bit."""
    from spinoza_amd import openqasm
    print(f"case qasm iqft text={by_text}", flush=True)
    if by_text:
        with open(_qasm("iqft.qasm")) as f:
            qc2 = openqasm.loads(f.read())
    else:
        qc2 = openqasm.load(_qasm("iqft.qasm"))
    ops = qc2.to_op_dicts()
    qc2.execute()
    n = qc2.num_qubits
    qc1 = spz.QuantumCircuit(spz.QuantumRegister(n))
    qc1.iqft(list(range(n))[::-1])
    qc1.execute()
    a, b = down(qc1.state), down(qc2.state)
    assert np.max(np.abs(a[0] - b[0])) < 1e-3 and np.max(np.abs(a[1] - b[1])) < 1e-3
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    assert exact(qc2.state, ref)


@pytest.mark.parametrize("by_text", [False, True])
def test_qasm_non_controlled_gates(spz, orc, by_text):
    """openqasm.rs:261-304 (non_controlled_gates_load / _loads)."""
    from spinoza_amd import openqasm
    print(f"case qasm test0 text={by_text}", flush=True)
    if by_text:
        with open(_qasm("test0.qasm")) as f:
            qc2 = openqasm.loads(f.read())
    else:
        qc2 = openqasm.load(_qasm("test0.qasm"))
    qc2.execute()
    n = qc2.num_qubits
    qc1 = spz.QuantumCircuit(spz.QuantumRegister(n))
    qc1.h(0); qc1.x(1); qc1.y(2); qc1.z(3); qc1.rx(1.0, 4); qc1.ry(2.0, 5); qc1.rz(3.0, 6)
    qc1.u(1.0, 2.0, 3.0, 7)
    ops = qc1.to_op_dicts()
    qc1.execute()
    a, b = down(qc1.state), down(qc2.state)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    assert exact(qc2.state, ref)


def test_qasm_quantum_lstm(spz, orc):
    """openqasm.rs:194-213 (quantum_lstm_load executes the program); here
    the result is also checked against the oracle and its norm."""
    from spinoza_amd import openqasm
    print("case qasm lstm", flush=True)
    qc = openqasm.load(_qasm("quantum_lstm.qasm"))
    ops = qc.to_op_dicts()
    qc.execute()
    ref = orc.OracleState(qc.num_qubits)
    orc.execute(ref, ops)
    assert exact(qc.state, ref)
    assert abs(qc.state.norm() - 1.0) < 1e-14


def test_threads_on_different_handles(spz, orc):
    """Calls on different handles from several threads share the library's
    scratch (reduction partials, tile programs, dense-block fragments) under
    its lock: every thread gets the results of its own state."""
    import threading
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import random_circuit
    print("case threads", flush=True)
    cases = []
    for k, n in enumerate((12, 13, 14, 15)):
        ops = random_circuit(n, 3, 900 + k)
        ref = orc.OracleState(n)
        orc.execute(ref, ops)
        cases.append((n, ops, ref.norm(), ref.prob0(n - 1), ref.ising_energy(0.5)))
    errors = []

    def worker(n, ops, nrm, p0, en, mode):
        try:
            s = spz.State(n)
            for _ in range(6):
                s.set_basis(0)
                run_ops(s, ops, mode=mode)
                got = (s.norm(), spz.core.prob0(s, n - 1), spz.ising_energy(s, 0.5))
                if mode == N.EXEC_DENSE5:
                    ok = all(abs(a - b) < 1e-12 for a, b in zip(got, (nrm, p0, en)))
                else:
                    ok = got == (nrm, p0, en) or all(abs(a - b) < 1e-13
                                                     for a, b in zip(got, (nrm, p0, en)))
                if not ok:
                    errors.append((n, mode, got, (nrm, p0, en)))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append((n, mode, repr(e)))

    threads = []
    for (n, ops, nrm, p0, en), mode in zip(cases, (N.EXEC_AUTO, N.EXEC_DENSE5, N.EXEC_STRICT,
                                                   N.EXEC_DENSE5)):
        threads.append(threading.Thread(target=worker, args=(n, ops, nrm, p0, en, mode)))
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors[:4]
