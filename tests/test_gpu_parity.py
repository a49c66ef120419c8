"""GPU parity: the gfx950 kernels against the CPU oracle, through the C ABI.

Bar: max |amplitude difference| <= 1e-12 (the tolerance the target states for
FP64); the plain path is also expected to be bit-identical to the oracle
(same operations in the same order, no FMA contraction), which is asserted
separately so a loss of bit-exactness is visible even inside the tolerance.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
KINDS = ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U"]


@pytest.fixture(scope="module")
def spz():
    import spinoza_amd
    return spinoza_amd


def to_gpu(spz, os_, precision="f64"):
    s = spz.State(os_.n, precision)
    s.upload(os_.re, os_.im)
    return s


def assert_match(prod, ref, exact=True, tol=TOL, what=""):
    re, im = prod.download()
    err = max(np.max(np.abs(re - ref.re)), np.max(np.abs(im - ref.im)))
    assert err <= tol, f"{what}: max abs error {err}"
    if exact:
        assert np.array_equal(re, ref.re) and np.array_equal(im, ref.im), \
            f"{what}: not bit-identical (max err {err})"


def gate(spz, kind, th=0.0, ph=0.0, la=0.0):
    G = spz.Gate
    return {"H": G.H, "X": G.X, "Y": G.Y, "Z": G.Z, "P": G.P(th), "RX": G.RX(th),
            "RY": G.RY(th), "RZ": G.RZ(th), "U": G.U(th, ph, la)}[kind]


@pytest.mark.parametrize("n", [1, 2, 9, 16])
def test_apply_every_kind_every_target(spz, orc, n):
    print(f"case apply n={n}", flush=True)
    rng = np.random.default_rng(n)
    base = orc.random_state(n, n)
    for kind in KINDS:
        for t in range(n):
            th, ph, la = rng.uniform(0, 2 * np.pi, 3)
            ref = base.copy()
            ref.apply(kind, t, th, ph, la)
            s = to_gpu(spz, base)
            spz.apply(gate(spz, kind, th, ph, la), s, t)
            assert_match(s, ref, what=f"{kind} t={t}")


@pytest.mark.parametrize("n", [2, 7, 12])
def test_c_apply_every_pair(spz, orc, n):
    print(f"case c_apply n={n}", flush=True)
    rng = np.random.default_rng(100 + n)
    base = orc.random_state(200 + n, n)
    for kind in KINDS:
        for c in range(n):
            for t in range(n):
                if c == t:
                    continue
                th, ph, la = rng.uniform(0, 2 * np.pi, 3)
                ref = base.copy()
                ref.c_apply(kind, c, t, th, ph, la)
                s = to_gpu(spz, base)
                spz.c_apply(gate(spz, kind, th, ph, la), s, c, t)
                assert_match(s, ref, what=f"C{kind} c={c} t={t}")


def test_multi_control_masks(spz, orc):
    print("case mc_apply", flush=True)
    n = 9
    rng = np.random.default_rng(7)
    base = orc.random_state(9, n)
    for trial in range(60):
        kind = KINDS[trial % len(KINDS)]
        qs = rng.permutation(n)
        t = int(qs[0])
        ones = sum(1 << int(q) for q in qs[1:1 + rng.integers(1, 4)])
        zeros = sum(1 << int(q) for q in qs[4:4 + rng.integers(0, 3)])
        th, ph, la = rng.uniform(0, 2 * np.pi, 3)
        ref = base.copy()
        ref.mc_apply(kind, ones, zeros, t, th, ph, la)
        s = to_gpu(spz, base)
        spz.mc_apply_masked(gate(spz, kind, th, ph, la), s, ones, zeros, t)
        assert_match(s, ref, what=f"mc {kind} t={t} ones={ones:b} zeros={zeros:b}")
    # cc_apply (Toffoli) against the oracle
    for c0, c1, t in [(0, 1, 2), (2, 0, 1), (5, 8, 0), (3, 7, 6)]:
        ref = base.copy()
        ref.mc_apply("X", (1 << c0) | (1 << c1), 0, t)
        s = to_gpu(spz, base)
        spz.cc_apply(spz.Gate.X, s, c0, c1, t)
        assert_match(s, ref, what="ccx")


def test_swap_every_pair(spz, orc):
    print("case swap", flush=True)
    n = 8
    base = orc.random_state(3, n)
    for a in range(n):
        for b in range(n):
            ref = base.copy()
            ref.swap(a, b)
            s = to_gpu(spz, base)
            spz.apply(spz.Gate.SWAP(a, b), s, 0)
            assert_match(s, ref, what=f"swap {a} {b}")


def test_dense_blocks(spz, orc):
    print("case transform_u", flush=True)
    from spinoza_amd import Unitary
    n = 9
    rng = np.random.default_rng(3)
    base = orc.random_state(33, n)
    for m in (1, 2, 3, 4, 5):
        a = rng.normal(size=(1 << m, 1 << m)) + 1j * rng.normal(size=(1 << m, 1 << m))
        q, _ = np.linalg.qr(a)
        u = Unitary.from_matrix(q)
        for t in range(0, n - m + 1, 2):
            ref = base.copy()
            ref.transform_u(q, t)
            s = to_gpu(spz, base)
            spz.transform_u(s, u, t)
            assert_match(s, ref, what=f"dense m={m} t={t}")
            c = (t + m) % n
            if not (t <= c < t + m):
                ref = base.copy()
                ref.transform_u(q, t, control=c)
                s = to_gpu(spz, base)
                spz.c_transform_u(s, u, c, t)
                assert_match(s, ref, what=f"cdense m={m} t={t} c={c}")


def test_reference_kats(spz, orc, kats):
    """The reference's own known-answer tests, through the product."""
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from tests.kat_ops import expected_vectors, kat_op_dicts
    for case in kats["gate_cases"]:
        print(f"case kat {case['name']}", flush=True)
        for mode in (N.EXEC_PLAIN, N.EXEC_AUTO):
            s = spz.State(case["n"])
            run_ops(s, kat_op_dicts(case), mode=mode)
            idx, re, im = expected_vectors(case)
            got_re, got_im = s.download()
            np.testing.assert_allclose(got_re[idx], re, atol=case["tol"], rtol=0, err_msg=case["name"])
            np.testing.assert_allclose(got_im[idx], im, atol=case["tol"], rtol=0, err_msg=case["name"])


def test_measure_known_state_kat(spz, kats):
    print("case measure kat", flush=True)
    kc = kats["measure_known_state"]
    v = kc["interleaved_re_im"]
    s = spz.State.from_arrays(np.array(v[0::2]), np.array(v[1::2]))
    for step in kc["steps"]:
        out = spz.measure_qubit(s, step["target"], step["reset"], step["forced"])
        assert out == step["forced"]
        re, im = s.download()
        idx = [a[0] for a in step["amps"]]
        np.testing.assert_allclose(re[idx], [a[1] for a in step["amps"]], atol=kc["tol"])
        np.testing.assert_allclose(im[idx], [a[2] for a in step["amps"]], atol=kc["tol"])


@pytest.mark.parametrize("n", [3, 13, 20])
def test_reductions_bit_exact(spz, orc, n):
    print(f"case reductions n={n}", flush=True)
    from spinoza_amd import core
    base = orc.random_state(50 + n, n)
    s = to_gpu(spz, base)
    assert s.norm() == base.norm()
    for t in sorted({0, 1, n // 2, n - 1}):
        assert core.prob0(s, t) == base.prob0(t)
        assert spz.qubit_expectation_value(s, t) == base.qubit_expectation_value(t)
        for obs in "xyz":
            assert spz.xyz_expectation_value(obs, s, [t]) == base.xyz_expectation_value(obs, [t])
    assert spz.expect_zz(s, 0, n - 1) == base.expect_zz(0, n - 1)
    assert spz.ising_energy(s, 1.0) == base.ising_energy(1.0)
    with pytest.raises(ValueError):
        spz.xyz_expectation_value("a", s, [0])


@pytest.mark.parametrize("n", [4, 14])
def test_measurement_outcomes_bit_exact(spz, orc, n):
    print(f"case measure n={n}", flush=True)
    base = orc.random_state(60 + n, n)
    us = orc.uniforms(61 + n, 3 * n)
    ref = base.copy()
    s = to_gpu(spz, base)
    k = 0
    for rnd in range(3):
        for t in range(n):
            u = float(us[k]); k += 1
            reset = bool((t + rnd) & 1)
            v_ref, _ = ref.measure_qubit(t, reset=reset, forced=-1, u=u)
            v = spz.measure_qubit(s, t, reset, u=u)
            assert v == v_ref
            assert_match(s, ref, what=f"collapse t={t}")


def test_qft20_config(spz, orc):
    """The 20-qubit QFT workload: Gaussian input, all 2^20 amplitudes within
    1e-12 of the oracle and 4096 samples bit-identical."""
    print("case qft20", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import qft
    from oracle import textbook as tb
    n = 20
    base = orc.gaussian_state(1, n)
    ref = base.copy()
    orc.execute(ref, qft(n))
    for mode in (N.EXEC_PLAIN, N.EXEC_AUTO):
        s = to_gpu(spz, base)
        run_ops(s, qft(n), mode=mode)
        assert_match(s, ref, exact=(mode == N.EXEC_PLAIN), what=f"qft20 mode={mode}")
    # closed form: the QFT is the DFT
    dft = tb.qft_matrix_dft(base.re + 1j * base.im)
    assert np.max(np.abs((ref.re + 1j * ref.im) - dft)) < 1e-12
    u = orc.uniforms(2, 4096)
    assert np.array_equal(spz.sample(s, u), ref.sample(u))


def test_random_circuit_vs_oracle(spz, orc):
    print("case random circuit 18q", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import random_circuit
    n = 18
    ops = random_circuit(n, 6, 11)
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    for mode in (N.EXEC_PLAIN, N.EXEC_STRICT, N.EXEC_AUTO):
        s = spz.State(n)
        run_ops(s, ops, mode=mode)
        assert_match(s, ref, exact=(mode != N.EXEC_AUTO), what=f"random mode={mode}")


def test_error_paths(spz):
    print("case errors", flush=True)
    s = spz.State(4)
    with pytest.raises(spz.SpinozaError):
        spz.apply(spz.Gate.H, s, 4)
    with pytest.raises(spz.SpinozaError):
        spz.c_apply(spz.Gate.X, s, 2, 2)
    with pytest.raises(spz.SpinozaError):
        spz.apply(spz.Gate.SWAP(0, 9), s, 0)
    with pytest.raises(ValueError):
        spz.State(0)
    with pytest.raises(spz.SpinozaError) as e:
        spz.measure_qubit(s, 0, True, 1)  # |0000>: outcome 1 has probability 0
    assert e.value.code == 6
    with pytest.raises(spz.SpinozaError) as e:
        spz.State(40)  # 16 TiB: allocation must fail cleanly
    assert e.value.code == 3
    s2 = spz.State(4)
    assert s2.norm() == 1.0


def test_fp32_storage_fidelity(spz, orc):
    print("case fp32", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import random_circuit
    n = 16
    ops = random_circuit(n, 8, 5)
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    for mode in (N.EXEC_PLAIN, N.EXEC_AUTO):
        s = spz.State(n, "f32")
        run_ops(s, ops, mode=mode)
        psi = s.to_numpy()
        nrm = s.norm()
        # state fidelity of the (slightly unnormalised) FP32 state
        fid = abs(np.vdot(ref.re + 1j * ref.im, psi)) ** 2 / nrm
        assert 1 - fid < 1e-9, fid
        assert abs(nrm - 1.0) < 1e-5
        assert abs(spz.expect_zz(s, 0, n - 1) - ref.expect_zz(0, n - 1)) < 1e-5


def test_inner_mixed_precision(spz, orc):
    """<a|b> between an FP32 and an FP64 state (either order) against the
    product of the downloaded vectors, and the FP64 x FP64 case bit-exact
    against the oracle's fixed-tree dot product."""
    print("case inner mixed precision", flush=True)
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import random_circuit
    n = 14
    ops = random_circuit(n, 4, 9)
    a = spz.State(n, "f32")
    b = spz.State(n, "f64")
    run_ops(a, ops)
    run_ops(b, ops)
    pa, pb = a.to_numpy(), b.to_numpy()
    ref = np.vdot(pa, pb)
    for x, y, r in ((a, b, ref), (b, a, np.conj(ref))):
        v = spz.inner(x, y)
        assert abs(v - r) < 1e-12, (v, r)
    ob = orc.OracleState(n)
    orc.execute(ob, ops)
    c = to_gpu(spz, ob)
    assert spz.inner(c, c).real == ob.norm()


def test_state_beyond_2_32_amplitudes(spz):
    """A 33-qubit FP32 state (2^33 amplitudes, 64 GiB): grids are capped at
    the dispatch limit and the kernels stride, so the fill, a pair gate on
    the top qubit (2^32 pairs), a t = 0 gate, a diagonal gate and the norm
    cover every amplitude."""
    print("case 33q fp32", flush=True)
    n = 33
    s = spz.State(n, "f32")
    spz.apply(spz.Gate.H, s, n - 1)
    h = np.float32(2.0 ** -0.5)
    assert s[0] == (float(h), 0.0) and s[1 << (n - 1)] == (float(h), 0.0)
    spz.apply(spz.Gate.X, s, 0)
    spz.apply(spz.Gate.Z, s, n - 1)
    assert s[0] == (0.0, 0.0) and s[1] == (float(h), 0.0)
    assert s[(1 << (n - 1)) | 1] == (-float(h), 0.0)
    assert abs(s.norm() - 2 * float(h) ** 2) < 1e-7
    re, im = s.download((1 << n) - 4096, 4096)
    assert not re.any() and not im.any()
    del s


def test_full_size_properties(spz):
    """At the 30-qubit workload size: H on every qubit gives the uniform
    state exactly, X on every qubit maps |0> to |2^n - 1>, norm stays 1."""
    print("case 30q properties", flush=True)
    n = 30
    s = spz.State(n)
    for q in range(n):
        spz.apply(spz.Gate.H, s, q)
    re, im = s.download(0, 4096)
    assert np.all(re == re[0]) and np.all(im == 0.0)
    assert abs(re[0] - 2.0 ** (-15)) < 1e-18
    assert abs(s.norm() - 1.0) < 1e-13
    s.set_basis(0)
    for q in range(n):
        spz.apply(spz.Gate.X, s, q)
    assert s[(1 << n) - 1] == (1.0, 0.0) and s[0] == (0.0, 0.0)
    del s


@pytest.mark.parametrize("n", [12, 15])
def test_dense5_mfma_blocks_vs_oracle(spz, orc, n):
    """Gates fused into dense 5-qubit blocks and applied with FP64 MFMA."""
    print(f"case dense5 n={n}", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import qft, random_circuit
    for ops in (random_circuit(n, 5, 21 + n), qft(n)):
        ref = orc.OracleState(n)
        orc.execute(ref, ops)
        s = spz.State(n)
        run_ops(s, ops, mode=N.EXEC_DENSE5)
        assert_match(s, ref, exact=False, what=f"dense5 n={n}")


def test_dense5_fp32_fidelity(spz, orc):
    print("case dense5 fp32", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import random_circuit
    n = 16
    ops = random_circuit(n, 6, 3)
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    s = spz.State(n, "f32")
    run_ops(s, ops, mode=N.EXEC_DENSE5)
    psi = s.to_numpy()
    nrm = s.norm()
    fid = abs(np.vdot(ref.re + 1j * ref.im, psi)) ** 2 / nrm
    assert 1 - fid < 1e-8, fid
    assert abs(nrm - 1.0) < 1e-5


def test_controlled_gates_in_tiles_vs_oracle(spz, orc):
    """Every gate kind with one or two controls (1- and 0-controls) on qubits
    that land on register bits, lane bits and outside the tile, at 16 qubits
    so the tile executor runs: STRICT bit-identical to the oracle, AUTO within
    1e-12 (the oracle applies the same masks, restated from the reference)."""
    print("case controlled gates in tiles", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import op
    n = 16
    rng = np.random.default_rng(77)
    kinds = [N.H, N.X, N.Y, N.Z, N.P, N.RX, N.RY, N.RZ, N.U]
    ops = []
    for i in range(400):
        k = kinds[i % len(kinds)]
        t = int(rng.integers(n))
        others = [q for q in range(n) if q != t]
        cs = rng.choice(others, size=int(rng.integers(0, 3)), replace=False)
        ones = zeros = 0
        for c in cs:
            if rng.random() < 0.75:
                ones |= 1 << int(c)
            else:
                zeros |= 1 << int(c)
        ops.append(op(k, t, theta=float(rng.uniform(0, 6.28)), phi=float(rng.uniform(0, 6.28)),
                      lam=float(rng.uniform(0, 6.28)), ones=ones, zeros=zeros))
        if i % 37 == 0:  # keep amplitudes spread out
            for q in range(n):
                ops.append(op(N.H, q))
    base = orc.random_state(5, n)
    ref = base.copy()
    orc.execute(ref, ops)
    for mode in (N.EXEC_STRICT, N.EXEC_AUTO):
        s = to_gpu(spz, base)
        run_ops(s, ops, mode=mode)
        assert_match(s, ref, exact=(mode == N.EXEC_STRICT), what=f"controlled tiles mode={mode}")


@pytest.mark.gpu
def test_back_to_back_executes_without_sync(spz, orc):
    """Execute calls return without a device synchronisation and the tile
    programs rotate through two staging slots: five different circuits queued
    back to back on five states (the third call reuses the first call's slot
    while earlier sweeps may still run) must each match the oracle."""
    print("case back-to-back executes", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import random_circuit
    n = 16
    cases = []
    for k in range(5):
        ops = random_circuit(n, 6 + 3 * k, 100 + k)
        base = orc.random_state(20 + k, n)
        ref = base.copy()
        orc.execute(ref, ops)
        cases.append((ops, base, ref))
    states = [to_gpu(spz, base) for _, base, _ in cases]
    N.check(N.lib().spz_sync())
    for s, (ops, _, _) in zip(states, cases):
        run_ops(s, ops, mode=N.EXEC_STRICT)
    for k, (s, (_, _, ref)) in enumerate(zip(states, cases)):
        assert_match(s, ref, exact=True, what=f"back-to-back circuit {k}")


@pytest.mark.gpu
@pytest.mark.parametrize("tile_qubits", [10, 12])
def test_lane_swaps_and_fast_sweeps_vs_oracle(spz, orc, tile_qubits, monkeypatch):
    """Sweeps whose layout changes include in-wavefront lane swaps (register
    bit <-> lane bit 4 or 5, v_permlane16/32_swap; off by default, turned on
    here with SPZ_TILE_SWAPS=1), and the AUTO arithmetic:
    runs of H with nothing to absorb their factor before the sweep ends, H
    followed by RX / RY / U / RZ that absorb it, controlled H (which applies
    its own factor).  STRICT (swaps, reference arithmetic) is bit-identical to
    the oracle; AUTO within 1e-12.  The plan must contain lane swaps."""
    print(f"case lane swaps / fast sweeps K={tile_qubits}", flush=True)
    monkeypatch.setenv("SPZ_TILE_SWAPS", "1")
    from spinoza_amd import _native as N
    from spinoza_amd.executor import plan_stats, run_ops
    from spinoza_amd.workloads import op
    n = 15
    rng = np.random.default_rng(91 + tile_qubits)
    ops = []
    for layer in range(12):
        for q in range(n):
            k = [N.H, N.H, N.RX, N.RY, N.U, N.RZ, N.Y, N.X][int(rng.integers(8))]
            ops.append(op(k, q, theta=float(rng.uniform(0, 6.28)), phi=float(rng.uniform(0, 6.28)),
                          lam=float(rng.uniform(0, 6.28))))
        for _ in range(4):
            c, t = (int(x) for x in rng.choice(n, 2, replace=False))
            ops.append(op(N.H if layer % 3 == 0 else N.X, t, ones=1 << c))
        for q in range(0, n, 3):  # runs of bare H
            ops.append(op(N.H, q))
            ops.append(op(N.H, (q + 1) % n))
    st = plan_stats(n, ops, mode=N.EXEC_AUTO, tile_qubits=tile_qubits)
    ss = plan_stats(n, ops, mode=N.EXEC_STRICT, tile_qubits=tile_qubits)
    assert st["lane_swaps"] > 0 and ss["lane_swaps"] > 0, (st, ss)
    base = orc.random_state(12, n)
    ref = base.copy()
    orc.execute(ref, ops)
    for mode in (N.EXEC_STRICT, N.EXEC_AUTO):
        s = to_gpu(spz, base)
        run_ops(s, ops, mode=mode, tile_qubits=tile_qubits)
        assert_match(s, ref, exact=(mode == N.EXEC_STRICT), what=f"swaps/fast mode={mode}")


@pytest.mark.gpu
@pytest.mark.parametrize("tile_qubits,precision", [(10, "f64"), (12, "f64"), (11, "f32")])
def test_pauli_frames_vs_oracle(spz, orc, tile_qubits, precision):
    """AUTO sweeps absorb uncontrolled X, Y and Z into a Pauli frame
    (psi = i^w X^x Z^z phi), carry it through H, CX and CZ, conjugate every
    other gate by it (RX/RY/RZ/U sign changes, P and controlled Z under X as
    general diagonals, U under X and controlled X/Y/H as general matrices,
    controls under X inverted), and apply it at the store (index XOR, sign
    parity, an imaginary phase as (-im, re)).  Circuits rich in Paulis, with
    CX/CZ whose control lies outside the tile (Z frame bits outside the
    tile), zero-controls, multi-controls and a Pauli-only tail.  FP64 within
    1e-12 of the oracle; FP32 storage within 1e-6."""
    print(f"case pauli frames K={tile_qubits} {precision}", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import plan_stats, run_ops
    from spinoza_amd.workloads import op
    n = 14
    rng = np.random.default_rng(7 + tile_qubits)
    single = [N.X, N.Y, N.Z, N.X, N.Y, N.Z, N.H, N.RX, N.RY, N.RZ, N.P, N.U]
    ops = []

    def ang():
        return float(rng.uniform(0, 6.28))

    for layer in range(10):
        for q in range(n):
            ops.append(op(single[int(rng.integers(len(single)))], q, theta=ang(), phi=ang(),
                          lam=ang()))
        perm = rng.permutation(n)
        for p in range(n // 2):
            c, t = int(perm[2 * p]), int(perm[2 * p + 1])
            ops.append(op(N.X if rng.integers(2) else N.Z, t, ones=1 << c))
        for _ in range(3):
            c1, c2, t = (int(x) for x in rng.choice(n, 3, replace=False))
            k = [N.U, N.Y, N.H, N.P, N.RX, N.X, N.Z][int(rng.integers(7))]
            ones, zeros = [(1 << c1, 1 << c2), (1 << c1 | 1 << c2, 0), (0, 1 << c1),
                           (1 << c1, 0)][int(rng.integers(4))]
            ops.append(op(k, t, theta=ang(), phi=ang(), lam=ang(), ones=ones, zeros=zeros))
    for _ in range(25):  # Pauli-only tail, an odd number of Y
        ops.append(op([N.X, N.Z][int(rng.integers(2))], int(rng.integers(n))))
    ops.append(op(N.Y, int(rng.integers(n))))
    st = plan_stats(n, ops, mode=N.EXEC_AUTO, tile_qubits=tile_qubits, precision=precision)
    assert st["frame_gates"] > 0 and st["tile_gates"] + st["plain_gates"] == len(ops), st
    base = orc.random_state(3 + tile_qubits, n)
    ref = base.copy()
    orc.execute(ref, ops)
    s = to_gpu(spz, base, precision)
    run_ops(s, ops, mode=N.EXEC_AUTO, tile_qubits=tile_qubits)
    assert_match(s, ref, exact=False, tol=TOL if precision == "f64" else 1e-6,
                 what=f"pauli frames K={tile_qubits} {precision}")


@pytest.mark.gpu
@pytest.mark.parametrize("tile_qubits", [10, 12])
def test_fast_rotation_edge_angles_vs_oracle(spz, orc, tile_qubits):
    """AUTO sweeps run RX/RY/RZ/P as rotations by three shears, which need
    |angle| <= pi/2: a larger angle runs as angle - pi with its sign carried
    (an uncontrolled RX/RY as a global -1, an uncontrolled P as a Z in the
    frame, an uncontrolled RZ or general diagonal as a phase on |1> with the
    |0> phase in the sweep's global factor), a controlled one as two ops at
    half the angle.  Angles at and around the switch points (0, +-pi/2,
    +-pi, pi +- 1e-9, 2 pi, 4 pi - 1e-9), uncontrolled, 1- and 0-controlled,
    on every qubit of a 14-qubit state (register, lane and outside-the-tile
    targets).  Within 1e-12 of the oracle."""
    print(f"case fast rotation edge angles K={tile_qubits}", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    from spinoza_amd.workloads import op
    n = 14
    pi = math.pi
    angles = [0.0, pi / 2, -pi / 2, pi / 2 + 1e-12, pi, -pi, pi - 1e-9, pi + 1e-9, 1.5 * pi,
              2 * pi, -2 * pi, 3 * pi, 4 * pi - 1e-9, 0.3, -2.9]
    rng = np.random.default_rng(17 + tile_qubits)
    ops = []
    for i, th in enumerate(angles * 3):
        for q in range(n):
            kind = [N.RX, N.RY, N.RZ, N.P][(i + q) % 4]
            c = (q + 1 + i % (n - 1)) % n  # never the target
            ctl = [dict(), dict(ones=1 << c), dict(zeros=1 << c)][(i // len(angles) + q) % 3]
            ops.append(op(kind, q, theta=th, **ctl))
        ops.append(op(N.H, int(rng.integers(n))))
        c, t = (int(x) for x in rng.choice(n, 2, replace=False))
        ops.append(op(N.X, t, ones=1 << c))
    base = orc.random_state(29 + tile_qubits, n)
    ref = base.copy()
    orc.execute(ref, ops)
    s = to_gpu(spz, base)
    run_ops(s, ops, mode=N.EXEC_AUTO, tile_qubits=tile_qubits)
    assert_match(s, ref, exact=False, what=f"edge angles K={tile_qubits}")


@pytest.mark.gpu
@pytest.mark.parametrize("tile_qubits", [10, 12])
def test_fast_general_2x2_vs_oracle(spz, orc, tile_qubits):
    """AUTO sweeps run an uncontrolled U, and any gate the Pauli frame turns
    into a general matrix, as a phase on |1>, an RY and a phase on |1>
    (three-shear rotations) with its global phase in the sweep's factor.
    The split takes cos and sin of the RY from |u00| and |u10| and its phases
    from the larger entries, then moves both phase angles into [-pi/2, pi/2]
    (a Z into the frame when needed).  U at the switch points: theta where
    sin or cos of theta/2 is 0 or tiny, phi and lambda at 0, +-pi/2, pi,
    3 pi/2 and near 0, each on a qubit that an X, Y or Z (absorbed into the
    frame) may precede, between H and CX/CZ that move the frame; controlled
    U keep the 2x2 product form.  Within 1e-12 of the oracle."""
    print(f"case fast general 2x2 K={tile_qubits}", flush=True)
    from spinoza_amd import _native as N
    from spinoza_amd.executor import plan_stats, run_ops
    from spinoza_amd.workloads import op
    n = 14
    pi = math.pi
    thetas = [0.0, pi, 2 * pi, 3 * pi, pi / 2, -pi / 2, 1e-9, pi - 1e-9, 0.7]
    phases = [0.0, pi / 2, -pi / 2, pi, 1.5 * pi, 1e-12, 2.5]
    rng = np.random.default_rng(41 + tile_qubits)
    ops = []
    i = 0
    for th in thetas:
        for ph in phases:
            for la in phases:
                q = i % n
                pre = [None, N.X, N.Y, N.Z][(i // n) % 4]
                if pre is not None:
                    ops.append(op(pre, q))
                if i % 11 == 5:
                    c = (q + 1 + i % (n - 1)) % n
                    ops.append(op(N.U, q, theta=th, phi=ph, lam=la, ones=1 << c))
                else:
                    ops.append(op(N.U, q, theta=th, phi=ph, lam=la))
                if i % 5 == 0:
                    ops.append(op(N.H, int(rng.integers(n))))
                if i % 3 == 0:
                    c, t = (int(x) for x in rng.choice(n, 2, replace=False))
                    ops.append(op(N.X if i % 2 else N.Z, t, ones=1 << c))
                i += 1
    st = plan_stats(n, ops, mode=N.EXEC_AUTO, tile_qubits=tile_qubits)
    assert st["frame_gates"] > 0, st
    base = orc.random_state(53 + tile_qubits, n)
    ref = base.copy()
    orc.execute(ref, ops)
    s = to_gpu(spz, base)
    run_ops(s, ops, mode=N.EXEC_AUTO, tile_qubits=tile_qubits)
    assert_match(s, ref, exact=False, what=f"general 2x2 K={tile_qubits}")
