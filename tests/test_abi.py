"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, the ctypes structs match the C layout, and calls fail
cleanly (no compute) when no device is visible."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spinoza_amd.h")
LIB = os.path.join(ROOT, "spinoza_amd", "libspinoza_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spz_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (spz_[a-z0-9_]+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    from spinoza_amd import _native
    names = declared_functions()
    assert set(names) == set(_native.SIGNATURES), set(names) ^ set(_native.SIGNATURES)
    lib = _native.load()
    for n in names:
        assert getattr(lib, n) is not None


def test_struct_layout_matches_c(tmp_path):
    from spinoza_amd import _native as N
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "spinoza_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(spz_op), sizeof(spz_gate), sizeof(spz_tracker), sizeof(spz_exec_opts));
  printf("%zu %zu %zu %zu %zu %zu %zu\n", offsetof(spz_op, theta), offsetof(spz_op, t1),
         offsetof(spz_op, ones), offsetof(spz_op, u), offsetof(spz_op, m), offsetof(spz_op, u_re),
         offsetof(spz_op, u_im));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(spz_batch_circuit), sizeof(spz_batch_opts),
         sizeof(spz_batch_stats), sizeof(spz_plan_info), offsetof(spz_batch_stats, n_small),
         offsetof(spz_plan_info, plain_gates));
  return 0;
}
''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = list(map(int, lines[0].split()))
    offs = list(map(int, lines[1].split()))
    assert sizes == [ctypes.sizeof(N.Op_t), ctypes.sizeof(N.Gate_t), ctypes.sizeof(N.Tracker_t),
                     ctypes.sizeof(N.ExecOpts_t)]
    assert offs == [N.Op_t.theta.offset, N.Op_t.t1.offset, N.Op_t.ones.offset, N.Op_t.u.offset,
                    N.Op_t.m.offset, N.Op_t.u_re.offset, N.Op_t.u_im.offset]
    more = list(map(int, lines[2].split()))
    assert more == [ctypes.sizeof(N.BatchCircuit_t), ctypes.sizeof(N.BatchOpts_t),
                    ctypes.sizeof(N.BatchStats_t), ctypes.sizeof(N.PlanInfo_t),
                    N.BatchStats_t.n_small.offset, N.PlanInfo_t.plain_gates.offset]


def test_enum_values_match_header():
    from spinoza_amd import _native as N
    text = open(HEADER).read()
    for name, val in [("SPZ_H", N.H), ("SPZ_U", N.U), ("SPZ_SWAP", N.SWAP), ("SPZ_M", N.M),
                      ("SPZ_UNITARY", N.UNITARY), ("SPZ_ERR_PROB", N.ERR_PROB),
                      ("SPZ_EXEC_STRICT", N.EXEC_STRICT), ("SPZ_K_TILE", N.K_TILE)]:
        m = re.search(rf"\b{name}\s*=\s*(\d+)", text)
        assert m and int(m.group(1)) == val, name


def test_oracle_codes_match_header():
    from oracle import pyoracle
    from spinoza_amd import _native as N
    assert (pyoracle.H, pyoracle.U, pyoracle.SWAP, pyoracle.M, pyoracle.UNITARY) == \
        (N.H, N.U, N.SWAP, N.M, N.UNITARY)


def test_no_device_fails_cleanly():
    """Without a GPU the first call reports SPZ_ERR_NODEVICE; nothing falls
    back to the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import spinoza_amd as spz
    with pytest.raises(spz.SpinozaError) as e:
        spz.State(3)
    assert e.value.code == 7


def test_gate_host_logic():
    """Gate::inverse / to_matrix host logic (gates.rs:78-190)."""
    import numpy as np
    from spinoza_amd import Gate
    for g in [Gate.H, Gate.X, Gate.Y, Gate.Z, Gate.P(2.03), Gate.RX(2.03), Gate.RY(3.03),
              Gate.RZ(3.03), Gate.U(1.0, 2.0, 3.0)]:
        m = g.to_matrix() @ g.inverse().to_matrix()
        assert np.max(np.abs(m - np.eye(2))) < 1e-15
        assert np.max(np.abs(g.to_matrix() @ g.to_matrix().conj().T - np.eye(2))) < 1e-15
    with pytest.raises(ValueError):
        Gate.M.inverse()
    with pytest.raises(ValueError):
        Gate.SWAP(0, 1).inverse().to_matrix()


def test_circuit_host_logic():
    """Register shifting, control bookkeeping and op encoding (no GPU)."""
    from spinoza_amd.circuit import Controls, QuantumRegister
    q0, q1 = QuantumRegister(3), QuantumRegister(2)
    bits = 0
    for r in (q0, q1):
        r.update_shift(bits)
        bits += len(r)
    assert q1.get_shift() == 3 and list(q1.qubits) == [3, 4]
    c = Controls.single(1).new_with_control(7, 3)
    assert c.kind == Controls.ONES and c.controls == [4, 7]
    c = Controls.none().new_with_control(2, 0)
    assert c.kind == Controls.SINGLE and c.controls == [2]
    c = Controls.mixed([0, 1], {1}).new_with_control(5, 2)
    assert c.kind == Controls.MIXED and c.controls == [2, 3, 5] and c.zeros == {3}


def test_workload_generators():
    from spinoza_amd import workloads as W
    ops = W.random_circuit(30, 40, 2)
    assert len(ops) == 40 * 45
    assert ops == W.random_circuit(30, 40, 2)
    for layer in range(40):
        pairs = ops[layer * 45 + 30: layer * 45 + 45]
        qs = [o["target"] for o in pairs] + [o["ones"].bit_length() - 1 for o in pairs]
        assert sorted(qs) == list(range(30))  # a perfect matching
    h, thetas, where = W.hea(24, 10, seed=4)
    assert len(thetas) == 480 and len(where) == 480
    assert len(h) == 10 * (48 + 23)
    q = W.qft(20)
    assert len(q) == 20 + 190 + 10


def test_splitmix_vectorised_matches_scalar():
    from spinoza_amd.rng import SplitMix64
    a, b = SplitMix64(42), SplitMix64(42)
    va = a.uniforms(1000)
    vb = [b.uniform() for _ in range(1000)]
    assert list(va) == vb and a.state == b.state
    from oracle import pyoracle
    assert list(pyoracle.uniforms(42, 1000)) == vb
