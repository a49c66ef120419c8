"""Correctness of the benchmark configurations at their checking sizes (slow:
each needs minutes of CPU time in the oracle).

* 30-qubit random circuit, checked on its 26-qubit replica: the benchmark's
  generator (seed 2) for all 40 layers, 1,800 gates, through the fused tile
  executor: STRICT bit-identical to the oracle, AUTO within 1e-12.
* 24-qubit hardware-efficient ansatz, 10 layers, 480 parameters (seed 4, the
  benchmark's adjoint workload): energy within 1e-12 of the oracle, a seeded
  32-parameter subset of the adjoint gradient within 1e-10 of oracle
  parameter shift (exact for RY/RZ up to rounding), and three parameters
  against central finite differences with h = 1e-6 (a sanity bound of 1e-7:
  the difference quotient carries ~eps |E| / h of rounding).

The oracle shares the forward prefix between shifted evaluations: for a
parameter at op i it runs ops [0, i) once, then the shifted op and the
suffix for each sign.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spz():
    import spinoza_amd
    return spinoza_amd


@pytest.fixture(scope="module")
def replica(orc):
    from spinoza_amd.workloads import random_circuit
    orc.set_threads(16)
    n = 26
    ops = random_circuit(n, 40, 2)
    print(f"case replica 26q oracle ({len(ops)} gates)", flush=True)
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    return n, ops, ref


@pytest.mark.timeout(600)
@pytest.mark.parametrize("mode", ["strict", "auto"])
def test_replica_26q_all_layers(spz, replica, mode):
    from spinoza_amd import _native as N
    from spinoza_amd.executor import run_ops
    n, ops, ref = replica
    print(f"case replica 26q {mode}", flush=True)
    assert len(ops) == 40 * (26 + 13)
    s = spz.State(n)
    run_ops(s, ops, mode=N.EXEC_STRICT if mode == "strict" else N.EXEC_AUTO)
    re, im = s.download()
    err = max(np.max(np.abs(re - ref.re)), np.max(np.abs(im - ref.im)))
    print(f"replica 26q {mode}: max abs error {err:.3e}", flush=True)
    assert err <= 1e-12
    if mode == "strict":
        assert np.array_equal(re, ref.re) and np.array_equal(im, ref.im)


@pytest.mark.timeout(300)
def test_endgame_plan_28q_auto_vs_plain(spz):
    """At 2^28 amplitudes and more the planner looks ahead at the end of a
    run so that no gate is left to a plain pass.  The 28-qubit, 60-layer
    random circuit of seed 2 is one where the greedy plan did leave one; its
    AUTO result against the PLAIN one (one kernel per gate, bit-identical to
    the oracle in the other tests; the oracle is too slow at this size):
    fidelity 1 within 1e-13, norms and <Z0 Z27> within 1e-12."""
    from spinoza_amd import _native as N
    from spinoza_amd.executor import plan_stats, run_ops
    from spinoza_amd.workloads import random_circuit
    n = 28
    ops = random_circuit(n, 60, 2)
    st = plan_stats(n, ops, mode=N.EXEC_AUTO)
    assert st["plain_gates"] == 0 and st["tile_gates"] == len(ops), st
    print(f"case endgame 28q ({len(ops)} gates, {st['sweeps']} sweeps)", flush=True)
    ref = spz.State(n)
    run_ops(ref, ops, mode=N.EXEC_PLAIN)
    s = spz.State(n)
    run_ops(s, ops, mode=N.EXEC_AUTO)
    nr, ns = ref.norm(), s.norm()
    fid = abs(spz.inner(s, ref)) ** 2 / (nr * ns)
    print(f"endgame 28q: infidelity {1 - fid:.3e}, norms {nr!r} {ns!r}", flush=True)
    assert abs(1.0 - fid) <= 1e-13
    assert abs(nr - ns) <= 1e-12
    assert abs(spz.expect_zz(s, 0, n - 1) - spz.expect_zz(ref, 0, n - 1)) <= 1e-12


def shifted_energies(orc, n, ops, idxs, shift, hx):
    """{i: (E(theta_i + shift), E(theta_i - shift))} for the ops idxs."""
    st = orc.OracleState(n)
    done = 0
    out = {}
    for i in sorted(idxs):
        orc.execute(st, ops[done:i])
        done = i
        vals = []
        for sgn in (1.0, -1.0):
            s2 = st.copy()
            o = dict(ops[i])
            o["theta"] = o["theta"] + sgn * shift
            orc.execute(s2, [o] + ops[i + 1:])
            vals.append(s2.ising_energy(hx))
            del s2
        out[i] = tuple(vals)
    return out


@pytest.mark.timeout(900)
def test_adjoint_24q_hea_vs_oracle(spz, orc):
    from spinoza_amd.autograd import adjoint_gradient
    from spinoza_amd.rng import SplitMix64
    from spinoza_amd.workloads import hea
    orc.set_threads(16)
    n, layers, hx = 24, 10, 1.0
    ops, thetas, where = hea(n, layers, seed=4)
    assert len(where) == 480
    print("case adjoint 24q: GPU", flush=True)
    s = spz.State(n)
    e, g, drift = adjoint_gradient(s, ops, hx, ckpt_every=71)
    assert drift < 1e-12
    print("case adjoint 24q: oracle energy", flush=True)
    ref = orc.OracleState(n)
    orc.execute(ref, ops)
    e_ref = ref.ising_energy(hx)
    del ref
    assert abs(e - e_ref) <= 1e-12, (e, e_ref)
    rng = SplitMix64(2424)
    pick = set()
    while len(pick) < 32:
        pick.add(where[rng.next_u64() % len(where)])
    print("case adjoint 24q: oracle parameter shift on 32 parameters", flush=True)
    ps = shifted_energies(orc, n, ops, pick, math.pi / 2, hx)
    worst = 0.0
    for i, (ep, em) in ps.items():
        worst = max(worst, abs(g[i] - 0.5 * (ep - em)))
    print(f"adjoint 24q vs parameter shift: max abs diff {worst:.3e}", flush=True)
    assert worst <= 1e-10
    fd_pick = sorted(pick)[-3:]
    h = 1e-6
    print("case adjoint 24q: oracle central differences", flush=True)
    fd = shifted_energies(orc, n, ops, fd_pick, h, hx)
    for i, (ep, em) in fd.items():
        assert abs(g[i] - (ep - em) / (2 * h)) <= 1e-7, (i, g[i], (ep - em) / (2 * h))
