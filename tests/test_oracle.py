"""Pin the CPU oracle (oracle/spz_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign, a swapped operand, a wrong pair map) fails it:
  * the reference's own known-answer vectors (tests/golden/reference_kats.json);
  * an independent dense numpy simulator (oracle/textbook.py) on random
    states for every gate, every target, controls and 0-controls;
  * closed forms (expectation values, QFT = DFT, Ising energy);
  * invariants (norm preservation, gate * inverse = identity, sampling
    distribution) and brute force on tiny inputs.
"""
import math

import numpy as np
import pytest

from oracle import textbook as tb
from tests.kat_ops import expected_vectors, kat_op_dicts

KINDS = ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U"]
KIND_ID = {k: i for i, k in enumerate(KINDS)}


def _vec(s):
    return s.re + 1j * s.im


def _tb_apply(psi, n, kind, target, th=0.0, ph=0.0, la=0.0, ones=(), zeros=()):
    return tb.apply(psi, n, tb.gate_matrix(kind, th, ph, la), target, ones, zeros)


# ---------------------------------------------------------------------------
# Reference known-answer tests
# ---------------------------------------------------------------------------
def test_oracle_reference_gate_kats(orc, kats):
    for case in kats["gate_cases"]:
        s = orc.OracleState(case["n"])
        orc.execute(s, kat_op_dicts(case))
        idx, re, im = expected_vectors(case)
        tol = case["tol"]
        np.testing.assert_allclose(s.re[idx], re, atol=tol, rtol=0, err_msg=case["name"])
        np.testing.assert_allclose(s.im[idx], im, atol=tol, rtol=0, err_msg=case["name"])


def test_oracle_measure_known_state(orc, kats):
    kc = kats["measure_known_state"]
    v = kc["interleaved_re_im"]
    s = orc.OracleState(kc["n"], np.array(v[0::2]), np.array(v[1::2]))
    for step in kc["steps"]:
        out, _ = s.measure_qubit(step["target"], step["reset"], step["forced"])
        assert out == step["forced"]
        idx = [a[0] for a in step["amps"]]
        np.testing.assert_allclose(s.re[idx], [a[1] for a in step["amps"]], atol=kc["tol"])
        np.testing.assert_allclose(s.im[idx], [a[2] for a in step["amps"]], atol=kc["tol"])


def test_oracle_xyz_expectation_kat(orc, kats):
    e = kats["expectation"]["xyz_exp_val"]
    s = orc.OracleState(1)
    s.apply("RX", 0, 0.54)
    s.apply("RY", 0, 0.12)
    z = s.xyz_expectation_value("z", [0])[0]
    assert abs(z - e["expected"]) < e["tol"]
    # closed form cos(0.54) cos(0.12), much tighter than the reference's 1e-4
    assert abs(z - math.cos(0.54) * math.cos(0.12)) < 1e-15
    assert abs(s.qubit_expectation_value(0) - z) < 1e-15


def test_oracle_dense_colvec_kat(orc, kats):
    for c in kats["matrix_cases"]["unitary_column_vector_multiplication"]["cases"]:
        s = orc.OracleState(1, np.array(c["v_re"]), np.array(c["v_im"]))
        u = (np.array(c["u_re"]) + 1j * np.array(c["u_im"])).reshape(2, 2)
        s.transform_u(u, 0)
        assert list(s.re) == c["out_re"] and list(s.im) == c["out_im"]


# ---------------------------------------------------------------------------
# Against the independent textbook simulator
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kind", KINDS)
def test_oracle_single_gate_vs_textbook(orc, kind):
    n = 6
    rng = np.random.default_rng(11 + KIND_ID[kind])
    for target in range(n):
        base = orc.random_state(100 + target, n)
        th, ph, la = rng.uniform(0, 2 * np.pi, 3)
        s = base.copy()
        s.apply(kind, target, th, ph, la)
        ref = _tb_apply(_vec(base), n, kind, target, th, ph, la)
        assert np.max(np.abs(_vec(s) - ref)) < 1e-14, (kind, target)


@pytest.mark.parametrize("kind", KINDS)
def test_oracle_controlled_vs_textbook(orc, kind):
    n = 5
    rng = np.random.default_rng(21 + KIND_ID[kind])
    for control in range(n):
        for target in range(n):
            if control == target:
                continue
            base = orc.random_state(7 * control + target, n)
            th, ph, la = rng.uniform(0, 2 * np.pi, 3)
            s = base.copy()
            s.c_apply(kind, control, target, th, ph, la)
            ref = _tb_apply(_vec(base), n, kind, target, th, ph, la, ones=(control,))
            assert np.max(np.abs(_vec(s) - ref)) < 1e-14, (kind, control, target)


@pytest.mark.parametrize("kind", KINDS)
def test_oracle_multicontrol_vs_textbook(orc, kind):
    """Correct masks, including configurations where the reference's
    skip-ahead loop (gates.rs:425-446) pairs the wrong amplitudes."""
    n = 6
    rng = np.random.default_rng(31 + KIND_ID[kind])
    for trial in range(12):
        qs = rng.permutation(n)
        target = int(qs[0])
        ones = [int(q) for q in qs[1:1 + rng.integers(1, 3)]]
        zeros = [int(q) for q in qs[3:3 + rng.integers(0, 2)]]
        base = orc.random_state(1000 + trial, n)
        th, ph, la = rng.uniform(0, 2 * np.pi, 3)
        s = base.copy()
        s.mc_apply(kind, sum(1 << q for q in ones), sum(1 << q for q in zeros), target, th, ph, la)
        ref = _tb_apply(_vec(base), n, kind, target, th, ph, la, ones=ones, zeros=zeros)
        assert np.max(np.abs(_vec(s) - ref)) < 1e-14, (kind, target, ones, zeros)


def test_oracle_mc_ry_quirk_is_rx(orc):
    """ry_mc_apply calls the RX arithmetic (gates.rs:1209-1229); the quirk
    switch reproduces that, the default is a true RY."""
    base = orc.random_state(5, 4)
    a, b = base.copy(), base.copy()
    a.mc_apply("RY", 0b0011, 0, 3, 0.7, ry_as_rx=True)
    b.mc_apply("RX", 0b0011, 0, 3, 0.7)
    assert np.array_equal(a.re, b.re) and np.array_equal(a.im, b.im)


def test_oracle_swap_matches_three_cx(orc):
    """swap_apply vs utils::swap (3 CX), the equivalence the reference tests
    (gates.rs:1872-1901) -- exact."""
    n = 9
    for (a, b) in [(0, 1), (0, 8), (3, 5), (7, 2)]:
        base = orc.random_state(77 + a + b, n)
        s1, s2 = base.copy(), base.copy()
        s1.swap(a, b)
        s2.c_apply("X", a, b)
        s2.c_apply("X", b, a)
        s2.c_apply("X", a, b)
        assert np.array_equal(s1.re, s2.re) and np.array_equal(s1.im, s2.im)
        ref = tb.swap(_vec(base), n, a, b)
        assert np.array_equal(_vec(s1), ref)


def test_oracle_dense_block_vs_textbook(orc):
    n = 7
    rng = np.random.default_rng(5)
    for m in (1, 2, 3):
        for t in range(0, n - m + 1):
            a = rng.normal(size=(1 << m, 1 << m)) + 1j * rng.normal(size=(1 << m, 1 << m))
            q, _ = np.linalg.qr(a)
            base = orc.random_state(m * 10 + t, n)
            s = base.copy()
            s.transform_u(q, t)
            ref = tb.apply_block(_vec(base), n, q, t)
            assert np.max(np.abs(_vec(s) - ref)) < 1e-14
            for c in range(n):
                if t <= c < t + m:
                    continue
                s = base.copy()
                s.transform_u(q, t, control=c)
                ref = tb.apply_block(_vec(base), n, q, t, control=c)
                assert np.max(np.abs(_vec(s) - ref)) < 1e-14


def test_oracle_gate_matrix_matches_textbook(orc):
    import ctypes
    for kind in KINDS:
        for th, ph, la in [(0.3, 1.1, -0.4), (2.0, 3.0, 1.0)]:
            m = np.zeros(8)
            orc.lib().orc_gate_matrix(KIND_ID[kind], th, ph, la,
                                      m.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            mm = (m[0::2] + 1j * m[1::2]).reshape(2, 2)
            assert np.max(np.abs(mm - tb.gate_matrix(kind, th, ph, la))) < 1e-15


def test_oracle_gate_times_inverse_is_identity(orc):
    """gate . inverse = I (gates.rs:1903-2045), on a state, to rounding."""
    inv = {"P": lambda t, p, l: (-t, p, l), "RX": lambda t, p, l: (-t, p, l),
           "RY": lambda t, p, l: (-t, p, l), "RZ": lambda t, p, l: (-t, p, l),
           "U": lambda t, p, l: (-t, -l, -p)}
    base = orc.random_state(9, 5)
    for kind in KINDS:
        s = base.copy()
        th, ph, la = 1.0, 2.0, 3.0
        s.apply(kind, 2, th, ph, la)
        ti, pi_, li = inv.get(kind, lambda t, p, l: (t, p, l))(th, ph, la)
        s.apply(kind, 2, ti, pi_, li)
        assert np.max(np.abs(_vec(s) - _vec(base))) < 1e-15


# ---------------------------------------------------------------------------
# Closed forms and invariants
# ---------------------------------------------------------------------------
def test_oracle_qft_is_dft(orc):
    from spinoza_amd.workloads import qft
    n = 8
    s = orc.gaussian_state(3, n)
    psi0 = _vec(s).copy()
    orc.execute(s, qft(n))
    assert np.max(np.abs(_vec(s) - tb.qft_matrix_dft(psi0))) < 1e-13


def test_oracle_expectations_closed_form(orc):
    n = 6
    base = orc.random_state(44, n)
    psi = _vec(base)
    for t in range(n):
        for obs, kind in (("x", "X"), ("y", "Y"), ("z", "Z")):
            ref = np.vdot(psi, _tb_apply(psi, n, kind, t)).real
            assert abs(base.xyz_expectation_value(obs, [t])[0] - ref) < 1e-15
        p0 = np.sum(np.abs(psi[(np.arange(1 << n) >> t) & 1 == 0]) ** 2)
        assert abs(base.prob0(t) - p0) < 1e-15
    for a in range(n):
        for b in range(n):
            ref = np.vdot(psi, _tb_apply(_tb_apply(psi, n, "Z", a), n, "Z", b)).real
            assert abs(base.expect_zz(a, b) - ref) < 1e-15
    for hx in (1.0, 0.37):
        ref = np.vdot(psi, tb.ising_matrix_apply(psi, n, hx)).real
        assert abs(base.ising_energy(hx) - ref) < 1e-14
        lr, li = base.ising_apply(hx)
        assert np.max(np.abs((lr + 1j * li) - tb.ising_matrix_apply(psi, n, hx))) < 1e-14


def test_oracle_reduction_is_compensated(orc):
    """The fixed tree with TwoSum compensation sums 1 + many tiny terms
    exactly where naive left-to-right summation loses them."""
    n = 14
    N = 1 << n
    re = np.full(N, math.sqrt(1e-17))
    re[0] = 1.0
    s = orc.OracleState(n, re, np.zeros(N))
    exact = 1.0 + (N - 1) * 1e-17
    assert abs(s.norm() - exact) <= 2e-16
    naive = 0.0
    for v in re * re:
        naive += v
    assert naive == 1.0  # the tiny terms vanish without compensation


def test_oracle_measure_forced_and_drawn(orc):
    n = 5
    base = orc.random_state(8, n)
    for t in range(n):
        p0 = base.prob0(t)
        for u, expect in ((0.0, 1), (0.999999, 0), (1.0 - p0 - 1e-12, 1), (1.0 - p0 + 1e-12, 0)):
            s = base.copy()
            v, pp = s.measure_qubit(t, reset=False, forced=-1, u=u)
            assert v == expect and pp == p0
            assert abs(s.norm() - 1.0) < 1e-14
        s = base.copy()
        v, _ = s.measure_qubit(t, reset=True, forced=1)
        # reset moved the kept |1> half to |0>
        assert s.prob0(t) == pytest.approx(1.0, abs=1e-14)
    z = orc.OracleState(3)
    with pytest.raises(ValueError):
        z.measure_qubit(0, True, forced=1)


def test_oracle_sampling_distribution(orc):
    from scipy.stats import chisquare
    n = 6
    s = orc.random_state(12, n)
    u = orc.uniforms(99, 200000)
    idx = s.sample(u)
    counts = np.bincount(idx.astype(np.int64), minlength=1 << n)
    p = s.re ** 2 + s.im ** 2
    res = chisquare(counts, p / p.sum() * len(u))
    assert res.pvalue > 1e-4
    # degenerate basis states (core.rs:271-291)
    for i in range(8):
        b = orc.OracleState(3)
        b.re[0] = 0.0
        b.re[i] = 1.0
        assert set(b.sample(orc.uniforms(i, 1000)).tolist()) == {i}


def test_oracle_sampling_inverse_cdf_brute_force(orc):
    """The rule picks the first index whose running sum exceeds u * total."""
    n = 13  # two blocks of 4096
    s = orc.random_state(3, n)
    w = s.re * s.re + s.im * s.im
    u = orc.uniforms(5, 500)
    idx = s.sample(u)
    cdf = np.cumsum(w)
    for k in range(len(u)):
        x = u[k] * cdf[-1]
        brute = int(np.searchsorted(cdf, x, side="right"))
        assert abs(int(idx[k]) - brute) <= 1  # rounding of the two sums may differ by one step


@pytest.mark.parametrize("n,threads", [(10, 1), (16, 4)])
def test_cpu_baseline_loops_match_checker(orc, n, threads):
    """The CPU baseline (the reference's per-kind chunk loops and insert-bit
    controlled loops, vectorised, OpenMP above 15 qubits) gives the checker's
    results bit for bit for every gate kind, every target and control pairs
    below and above the target; CZ runs as c_apply(P(pi)) and differs from
    the exact negation only by sin(pi)."""
    rng = np.random.default_rng(n)
    ops = []
    kinds = ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U"]
    for t in range(n):
        for k in kinds:
            ops.append({"kind": k, "target": t, "theta": rng.uniform(0, 6.3),
                        "phi": rng.uniform(0, 6.3), "lam": rng.uniform(0, 6.3)})
    for _ in range(40):
        c, t = rng.choice(n, 2, replace=False)
        k = kinds[rng.integers(len(kinds))]
        if k == "Z":
            k = "X"
        ops.append({"kind": k, "target": int(t), "ones": 1 << int(c),
                    "theta": rng.uniform(0, 6.3), "phi": rng.uniform(0, 6.3),
                    "lam": rng.uniform(0, 6.3)})
    orc.set_threads(threads)
    try:
        a = orc.random_state(5, n)
        b = a.copy()
        orc.execute(a, ops)
        orc.rayon_execute(b, ops)
        assert np.array_equal(a.re, b.re) and np.array_equal(a.im, b.im)
        cz = [{"kind": "Z", "target": 3, "ones": 1 << 7}, {"kind": "Z", "target": 8, "ones": 1 << 1}]
        orc.execute(a, cz)
        orc.rayon_execute(b, cz)
        err = max(np.max(np.abs(a.re - b.re)), np.max(np.abs(a.im - b.im)))
        assert 0 < err <= 1e-15
        with pytest.raises(ValueError):
            orc.rayon_execute(b, [{"kind": "X", "target": 0, "ones": 6}])
    finally:
        orc.set_threads(1)
