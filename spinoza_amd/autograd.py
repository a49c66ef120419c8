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

    @classmethod
    def apply(cls, thetas, builder, n, hx=1.0, ckpt_every=0, precision="f64"):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx, th, builder_, n_, hx_, ck_, prec_):
                    vals = th.detach().to("cpu", torch.float64).tolist()
                    ops, param_ops = builder_(vals)
                    s = State(n_, prec_)
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
