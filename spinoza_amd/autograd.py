"""Gradients of the Ising-chain energy E(theta) = <psi(theta)|H|psi(theta)>.

H = sum_{i<n-1} Z_i Z_{i+1} + hx * sum_i X_i (open boundary).

* ``adjoint_gradient``  -- one forward pass plus the fused reverse sweep of the
  gfx950 library (``spz_adjoint_grad``): energy and dE/dtheta for every
  RX/RY/RZ/P op (controlled or not).
* ``param_shift_gradient`` -- the same gradient from 2 energy evaluations per
  parameter at theta +- pi/2 (exact for RX, RY, RZ and P), on the GPU.
* ``IsingEnergy`` -- a ``torch.autograd.Function`` whose backward returns the
  adjoint gradient, so the circuit drops into a torch optimisation loop.
"""
import ctypes
import math

import numpy as np

from . import _native as N
from .core import State, ising_energy
from .executor import encode_ops, run_ops

PARAM_KINDS = (N.RX, N.RY, N.RZ, N.P)


def adjoint_gradient(state, ops, hx=1.0, ckpt_every=0):
    """Run `ops` on `state` (left holding the final state) and return
    (energy, grad_per_op, drift); grad_per_op[i] is dE/dtheta of op i (0 for
    ops without a parameter)."""
    arr, keep = encode_ops(ops)
    grad = np.zeros(max(len(ops), 1))
    e = ctypes.c_double()
    drift = ctypes.c_double()
    N.check(N.lib().spz_adjoint_grad(state._h, arr, len(ops), float(hx), int(ckpt_every),
                                     N.dptr(grad), ctypes.byref(e), ctypes.byref(drift)))
    del keep
    return e.value, grad[:len(ops)], drift.value


def energy(n, ops, hx=1.0, precision="f64", state=None, mode=N.EXEC_AUTO):
    s = state if state is not None else State(n, precision)
    s.set_basis(0)
    run_ops(s, ops, mode=mode)
    return ising_energy(s, hx)


def param_shift_gradient(n, ops, param_ops, hx=1.0, subset=None, precision="f64",
                         mode=N.EXEC_AUTO):
    """dE/dtheta_k = (E(theta_k + pi/2) - E(theta_k - pi/2)) / 2 for the ops
    listed in `param_ops` (restricted to `subset` indices if given)."""
    s = State(n, precision)
    idx = range(len(param_ops)) if subset is None else subset
    out = {}
    for k in idx:
        i = param_ops[k]
        o = dict(ops[i])
        shifted = list(ops)
        o_p = dict(o, theta=o["theta"] + math.pi / 2)
        o_m = dict(o, theta=o["theta"] - math.pi / 2)
        shifted[i] = o_p
        ep = energy(n, shifted, hx, state=s, mode=mode)
        shifted[i] = o_m
        em = energy(n, shifted, hx, state=s, mode=mode)
        out[k] = 0.5 * (ep - em)
    return out


def _torch():
    import torch
    return torch


class IsingEnergy:
    """``IsingEnergy.apply(thetas, builder, n, hx, ckpt_every)``: `builder`
    maps a list of floats to (ops, param_ops); returns E as a 0-d float64
    tensor whose gradient w.r.t. `thetas` is the adjoint gradient."""

    _fn = None
    # One working state per (qubits, precision), reset to |0...0> on every
    # call: an optimiser loop then allocates (and frees, which synchronises
    # the device) nothing after its first step.
    _states = {}

    @classmethod
    def _state(cls, n, precision):
        s = cls._states.get((n, precision))
        if s is None:
            s = State(n, precision)
            cls._states[(n, precision)] = s
        else:
            s.set_basis(0)
        return s

    @classmethod
    def apply(cls, thetas, builder, n, hx=1.0, ckpt_every=0, precision="f64"):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, th, builder_, n_, hx_, ck_, prec_):
                    vals = th.detach().to("cpu", torch.float64).tolist()
                    ops, param_ops = builder_(vals)
                    s = cls._state(n_, prec_)
                    e, g, _ = adjoint_gradient(s, ops, hx_, ck_)
                    gt = torch.tensor([g[i] for i in param_ops], dtype=torch.float64)
                    ctx.save_for_backward(gt.to(th.device))
                    return torch.tensor(e, dtype=torch.float64, device=th.device)

                @staticmethod
                def backward(ctx, gout):
                    (gt,) = ctx.saved_tensors
                    return gout * gt, None, None, None, None, None

            cls._fn = _F
        return cls._fn.apply(thetas, builder, n, hx, ckpt_every, precision)


def batch_adjoint(circuits, hx=1.0, wg_max_qubits=0, use_graph=True, group_qubits=0,
                  wg_per_cu=0):
    """Energies and adjoint gradients of many independent circuits.

    `circuits` is a list of (n, ops); every circuit starts in |0...0>.
    Circuits up to `wg_max_qubits` run in the persistent kernel, those above
    `group_qubits` of them on a group of workgroups (0: the defaults; a
    negative group_qubits keeps every circuit in one workgroup), with
    `wg_per_cu` resident workgroups per CU (0: 2; negative: as many as fit).
    Returns (energies, [grad_per_op arrays], stats dict)."""
    all_ops, table, off, goff = [], [], 0, 0
    for n, ops in circuits:
        table.append((n, len(ops), off, goff))
        all_ops += ops
        off += len(ops)
        goff += len(ops)
    arr, keep = encode_ops(all_ops)
    cb = (N.BatchCircuit_t * max(len(table), 1))()
    for i, (n, k, o, g) in enumerate(table):
        cb[i].n, cb[i].n_ops, cb[i].op_offset, cb[i].grad_offset = n, k, o, g
    energies = np.zeros(max(len(table), 1))
    grads = np.zeros(max(goff, 1))
    opts = N.BatchOpts_t()
    opts.wg_max_qubits = int(wg_max_qubits)
    opts.use_graph = int(bool(use_graph))
    opts.group_qubits = int(group_qubits)
    opts.wg_per_cu = int(wg_per_cu)
    st = N.BatchStats_t()
    N.check(N.lib().spz_batch_adjoint(cb, len(table), arr, float(hx), N.dptr(energies),
                                      N.dptr(grads), ctypes.byref(opts), ctypes.byref(st)))
    del keep
    out = [grads[g:g + k].copy() for (n, k, o, g) in table]
    stats = {"total_ms": st.total_ms, "p50_ms": st.p50_ms, "p99_ms": st.p99_ms,
             "n_small": st.n_small, "n_large": st.n_large, "n_group": st.n_group,
             "wg_max_qubits": st.wg_max_qubits}
    return energies[:len(table)], out, stats


def batch_autotune(circuits):
    """Measure and keep the qubit count up to which circuits run one per
    workgroup (see spz_batch_autotune)."""
    all_ops, table, off = [], [], 0
    for n, ops in circuits:
        table.append((n, len(ops), off))
        all_ops += ops
        off += len(ops)
    arr, keep = encode_ops(all_ops)
    cb = (N.BatchCircuit_t * max(len(table), 1))()
    for i, (n, k, o) in enumerate(table):
        cb[i].n, cb[i].n_ops, cb[i].op_offset, cb[i].grad_offset = n, k, o, o
    out = ctypes.c_int()
    N.check(N.lib().spz_batch_autotune(cb, len(table), arr, ctypes.byref(out)))
    del keep
    return out.value
