"""Host-side tests of the tile scheduler (no GPU needed: spz_plan_stats only
plans).  Every gate of a barrier-free circuit must land in exactly one sweep or
one plain pass, the tile-set search must never need more sweeps than the plain
greedy choice, and program order (STRICT) can only need more sweeps."""
import os

import pytest

from spinoza_amd import _native as N
from spinoza_amd.executor import plan_stats
from spinoza_amd.workloads import hea, qft, random_circuit


def _with_search(passes, fn):
    old = os.environ.get("SPZ_TILE_SEARCH")
    os.environ["SPZ_TILE_SEARCH"] = str(passes)
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["SPZ_TILE_SEARCH"]
        else:
            os.environ["SPZ_TILE_SEARCH"] = old


@pytest.mark.parametrize("n,ops", [
    (30, random_circuit(30, 40, 2)),
    (26, random_circuit(26, 12, 7)),
    (20, qft(20)),
    (24, hea(24, 10, seed=4)[0]),
])
def test_every_gate_planned_once(n, ops):
    for mode in (N.EXEC_AUTO, N.EXEC_STRICT):
        st = plan_stats(n, ops, mode=mode)
        assert st["tile_gates"] + st["plain_gates"] == len(ops), (mode, st)
        assert st["sweeps"] >= 1


def test_tile_set_search_never_worse():
    for n, layers, seed in ((30, 40, 2), (28, 20, 5), (22, 30, 9)):
        ops = random_circuit(n, layers, seed)
        greedy = _with_search(0, lambda: plan_stats(n, ops))
        searched = plan_stats(n, ops)
        strict = plan_stats(n, ops, mode=N.EXEC_STRICT)
        assert searched["sweeps"] <= greedy["sweeps"], (n, greedy, searched)
        assert strict["sweeps"] >= searched["sweeps"]


def test_bench_circuit_sweep_count():
    """The 30-qubit, 40-layer bench circuit (1800 gates) fuses into at most 32
    sweeps of 12-qubit tiles (49 with the first-come tile set), and the
    end-of-run lookahead leaves none of its gates to a plain pass (the greedy
    tile sets left its last gate to one)."""
    st = plan_stats(30, random_circuit(30, 40, 2), tile_qubits=12)
    assert st["sweeps"] <= 32, st
    assert st["plain_gates"] == 0, st


@pytest.mark.parametrize("n,seed", [(28, 2), (30, 3), (32, 5)])
def test_endgame_lookahead_costs_no_sweeps(n, seed):
    """Random circuits whose greedy plan ends in a plain gate (28 q seed 2,
    32 q seed 5) or not (30 q seed 3): every gate still runs exactly once and
    the sweep counts stay those of the greedy plan (44, 32, 53)."""
    st = plan_stats(n, random_circuit(n, 40 if n == 30 else 60, seed), tile_qubits=12)
    assert st["tile_gates"] + st["plain_gates"] == (40 if n == 30 else 60) * (n + n // 2), st
    assert st["plain_gates"] == 0, st
    assert st["sweeps"] <= {28: 44, 30: 32, 32: 53}[n], st


def test_bench_circuit_layout_changes():
    """The lookahead register-set choice keeps the LDS layout changes of the
    bench circuit to about 3.3 per sweep (147 with the best-score choice)."""
    st = plan_stats(30, random_circuit(30, 40, 2), tile_qubits=12)
    assert st["layouts"] <= 110, st
    strict = plan_stats(30, random_circuit(30, 40, 2), tile_qubits=12, mode=N.EXEC_STRICT)
    assert strict["tile_gates"] + strict["plain_gates"] == 1800, strict


def test_plan_stats_rejects_bad_input():
    with pytest.raises(N.SpinozaError):
        plan_stats(0, [])
    with pytest.raises(N.SpinozaError):
        plan_stats(30, random_circuit(30, 1, 1), mode=N.EXEC_PLAIN)
