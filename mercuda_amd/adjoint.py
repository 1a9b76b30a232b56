"""Gradients of MVS integrations for orbit fitting: a thin
``torch.autograd.Function`` over libmercury_hip's tape
(mercury_hip_adjoint_forward / _backward / mercury_hip_tangent).

The integration itself never runs in torch: forward uploads the initial
states and masses, runs nsteps MVS steps on the device and keeps a tape
(every step's planet table, particle checkpoints every ~sqrt(nsteps) steps);
backward runs the hand-written adjoint over that tape.  Tensors are float64
on the CPU (the library's boundary takes host arrays).

    x0, v0, m = (torch tensors, requires_grad=True)
    xT, vT = MVSFlow.apply(x0, v0, m, nbig, nsteps, h)
    loss = ((xT[nbig:] - target) ** 2).sum()
    loss.backward()        # x0.grad, v0.grad, m.grad

Every final state may carry gradient, the Big bodies' as well as the
particles' (the central body's heliocentric state is identically zero);
masses receive gradient for Big bodies.
"""
from __future__ import annotations

import numpy as np
import torch

from .hip import MercuryHip

_LIB: MercuryHip | None = None


def library() -> MercuryHip:
    global _LIB
    if _LIB is None:
        _LIB = MercuryHip()
    return _LIB


class MVSFlow(torch.autograd.Function):
    """(x0[n,3], v0[n,3], m[n]) -> (x_T[n,3], v_T[n,3]) after nsteps MVS steps."""

    @staticmethod
    def forward(ctx, x0, v0, m, nbig, nsteps, h, segment=0, rce=None, rphys=None):
        lib = library()
        n = m.shape[0]
        lib.configure(1)
        lib.upload(m.detach().cpu().numpy(), x0.detach().cpu().numpy(), v0.detach().cpu().numpy(), int(nbig),
                   rce=np.zeros(n) if rce is None else rce, rphys=np.zeros(n) if rphys is None else rphys)
        xT, vT = lib.adjoint_forward(int(nsteps), float(h), int(segment))
        ctx.tape_id = _next_tape()
        ctx.nbig = int(nbig)
        return torch.from_numpy(xT), torch.from_numpy(vT)

    @staticmethod
    def backward(ctx, gx, gv):
        if ctx.tape_id != _TAPE[0]:
            raise RuntimeError("the library tape belongs to a later MVSFlow forward; "
                               "call backward before the next forward")
        lib = library()
        gx = np.zeros((lib.n, 3)) if gx is None else gx.detach().cpu().numpy().copy()
        gv = np.zeros((lib.n, 3)) if gv is None else gv.detach().cpu().numpy().copy()
        gx0, gv0, gm = lib.adjoint_backward(gx, gv)
        return (torch.from_numpy(gx0), torch.from_numpy(gv0), torch.from_numpy(gm),
                None, None, None, None, None, None)


_TAPE = [0]


def _next_tape() -> int:
    _TAPE[0] += 1
    return _TAPE[0]


def tangent(dx0, dv0, dm):
    """Tangent-linear pass over the tape of the last MVSFlow forward:
    derivative of (x_T, v_T) along (dx0, dv0, dm)."""
    lib = library()
    dx, dv = lib.tangent(np.asarray(dx0, float), np.asarray(dv0, float), np.asarray(dm, float))
    return torch.from_numpy(dx), torch.from_numpy(dv)
