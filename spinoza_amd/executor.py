"""Run plain op lists (dicts) through the C ABI executor ``spz_execute``.

An op dict has the keys of the C ABI's ``spz_op``: kind, target, theta, phi,
lam, t1, ones, zeros, classical (bool), u (uniform draw for M / BitFlipNoise)
and, for dense blocks, matrix (a 2^m x 2^m complex array).
"""
import ctypes

import numpy as np

from . import _native as N


def encode_ops(ops):
    n = len(ops)
    arr = (N.Op_t * max(n, 1))()
    keep = []
    for i, d in enumerate(ops):
        o = arr[i]
        o.kind = int(d["kind"])
        o.target = int(d.get("target", 0))
        o.theta = float(d.get("theta", 0.0))
        o.phi = float(d.get("phi", 0.0))
        o.lam = float(d.get("lam", 0.0))
        o.t1 = int(d.get("t1", 0))
        o.flags = N.OPF_CLASSICAL if d.get("classical") else 0
        o.ones = int(d.get("ones", 0))
        o.zeros = int(d.get("zeros", 0))
        o.u = float(d.get("u", 1.0))
        if "matrix" in d:
            u = np.asarray(d["matrix"], dtype=complex)
            re = np.ascontiguousarray(u.real)
            im = np.ascontiguousarray(u.imag)
            keep += [re, im]
            o.m = int(u.shape[0]).bit_length() - 1
            o.u_re = re.ctypes.data
            o.u_im = im.ctypes.data
    return arr, keep


def run_ops(state, ops, mode=N.EXEC_AUTO, tracker=None, tile_qubits=0, encoded=None):
    """Execute `ops` on `state`; returns the per-op outcome list (255 when an
    op is not a fresh measurement)."""
    arr, keep = encoded if encoded is not None else encode_ops(ops)
    n = len(ops)
    opts = N.ExecOpts_t()
    opts.mode = mode
    opts.tile_qubits = tile_qubits
    outcomes = (ctypes.c_uint8 * max(n, 1))()
    tr = tracker if tracker is not None else N.Tracker_t(0, 0)
    N.check(N.lib().spz_execute(state._h, arr, n, ctypes.byref(tr), ctypes.byref(opts),
                                outcomes))
    del keep
    return [int(outcomes[i]) for i in range(n)]


def plan_stats(n, ops, mode=N.EXEC_AUTO, tile_qubits=0, precision="f64"):
    """How the tile executor splits `ops` for an n-qubit state (host only, no
    device work): dict with sweeps, tile_gates, layouts (LDS transposes),
    lane_swaps (in-wavefront layout changes), frame_gates (X, Y, Z absorbed
    into a sweep's Pauli frame, counted in tile_gates too) and plain_gates."""
    arr, keep = encode_ops(ops)
    opts = N.ExecOpts_t()
    opts.mode = mode
    opts.tile_qubits = tile_qubits
    info = N.PlanInfo_t()
    prec = N.F64 if precision == "f64" else N.F32
    N.check(N.lib().spz_plan_stats(int(n), prec, arr, len(ops), ctypes.byref(opts),
                                   ctypes.byref(info)))
    del keep
    return {"sweeps": info.sweeps, "tile_gates": info.tile_gates, "layouts": info.layouts,
            "lane_swaps": info.lane_swaps, "frame_gates": info.frame_gates,
            "plain_gates": info.plain_gates}
