"""Adjoint of the dense 3D SL step and its torch.autograd wrapper.

``sl_adjoint`` calls ``dcl_sl_adjoint`` (the hand-written transposed gather,
adjoint.hip).  ``AdvectN`` is a torch.autograd.Function over N SL steps with
a steady velocity: the forward keeps one checkpoint of phi every
``checkpoint_every`` steps; the backward recomputes each segment from its
checkpoint and runs the adjoint steps in reverse, accumulating dL/dvel.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from .levelset import (BACKTRACE, INTERP, GridWindow, _dev, _p, advect_sl, precision_of,
                       stream_handle, velocity)


def sl_adjoint(phi0, vel, lam_next, win: GridWindow, dt: float, interp="cubic", backtrace="rk2",
               lam=None, gvel=None):
    """lam += dL/dphi0 and gvel += dL/dvel for one step with lam_next = dL/dphi."""
    _dev(phi0, "phi0")
    precision_of(phi0)
    _dev(vel, "vel", phi0.dtype)
    _dev(lam_next, "lam_next", phi0.dtype)
    for t, name in ((phi0, "phi0"), (lam_next, "lam_next")):
        if t.numel() != win.nodes:
            raise ValueError(f"{name} has {t.numel()} elements, window needs {win.nodes}")
    lam = torch.zeros_like(phi0) if lam is None else lam
    gvel = torch.zeros_like(vel) if gvel is None else gvel
    _dev(lam, "lam", phi0.dtype)
    _dev(gvel, "gvel", phi0.dtype)
    if lam.numel() != win.nodes:
        raise ValueError(f"lam has {lam.numel()} elements, window needs {win.nodes}")
    if gvel.shape != vel.shape:
        raise ValueError(f"gvel is shaped {tuple(gvel.shape)}, vel {tuple(vel.shape)}")
    V = velocity(vel, win)
    L.call("dcl_sl_adjoint", _p(phi0), C.byref(win.c()), C.byref(V), _p(lam_next), _p(lam), _p(gvel),
           float(dt), precision_of(phi0), INTERP[interp], BACKTRACE[backtrace], stream_handle())
    return lam, gvel


def adjoint_record(vel, win: GridWindow, dt: float, interp="cubic", backtrace="rk2"):
    """The departure record of a steady velocity (dcl_sl_adjoint_prepare): every
    interior node's stencil base, fractions and clamped axes, 48 bytes per node."""
    precision_of(vel)
    nbytes = C.c_int64()
    L.call("dcl_sl_adjoint_record_bytes", C.byref(win.c()), C.byref(nbytes))
    rec = torch.empty(nbytes.value, dtype=torch.uint8, device=vel.device)
    V = velocity(vel, win)
    L.call("dcl_sl_adjoint_prepare", C.byref(win.c()), C.byref(V), float(dt), precision_of(vel), INTERP[interp],
           BACKTRACE[backtrace], _p(rec), stream_handle())
    return rec


def adjoint_step(phi0, rec, lam_next, win: GridWindow, dt: float, lam, hsum, interp="cubic"):
    """lam += dL/dphi0 of one step from the record; hsum (3 per node) += the
    step's departure sensitivities (dcl_sl_adjoint_step)."""
    _dev(phi0, "phi0")
    prec = precision_of(phi0)
    for t, name, count in ((lam_next, "lam_next", 1), (lam, "lam", 1), (hsum, "hsum", 3)):
        _dev(t, name, phi0.dtype)
        if t.numel() != win.nodes * count:
            raise ValueError(f"{name} has {t.numel()} elements, the window needs {win.nodes * count}")
    if phi0.numel() != win.nodes:
        raise ValueError(f"phi0 has {phi0.numel()} elements, the window needs {win.nodes}")
    _dev(rec, "record", torch.uint8)
    if rec.numel() != 48 * win.nodes:
        raise ValueError("record does not belong to this window")
    L.call("dcl_sl_adjoint_step", _p(phi0), C.byref(win.c()), _p(rec), _p(lam_next), _p(lam), _p(hsum),
           float(dt), prec, INTERP[interp], stream_handle())
    return lam, hsum


def adjoint_finish(vel, hsum, win: GridWindow, dt: float, gvel, backtrace="rk2"):
    """gvel += dL/dvel formed from the accumulated sensitivities (dcl_sl_adjoint_finish)."""
    prec = precision_of(vel)
    _dev(hsum, "hsum", vel.dtype)
    _dev(gvel, "gvel", vel.dtype)
    if hsum.numel() != 3 * win.nodes or gvel.shape != vel.shape:
        raise ValueError("hsum or gvel does not match the window / velocity")
    V = velocity(vel, win)
    L.call("dcl_sl_adjoint_finish", C.byref(win.c()), C.byref(V), _p(hsum), _p(gvel), float(dt), prec,
           BACKTRACE[backtrace], stream_handle())
    return gvel


class AdvectN(torch.autograd.Function):
    """phi_N = SL^N(phi_0; vel) with checkpointed adjoint backward.  The
    velocity is the same at every step, so the backward computes the
    departure record once, runs one record-driven adjoint step per SL step
    (reverse order, states recomputed per checkpoint segment) and forms
    dL/dvel once from the accumulated sensitivities."""

    @staticmethod
    def forward(ctx, phi0, vel, win, dt, steps, interp, backtrace, checkpoint_every):
        ckpts = []
        phi = phi0.detach()
        for s in range(steps):
            if s % checkpoint_every == 0:
                ckpts.append(phi)
            phi = advect_sl(phi, vel.detach(), win, dt, interp, backtrace)
        ctx.save_for_backward(vel)
        ctx.ckpts = ckpts
        ctx.cfg = (win, dt, steps, interp, backtrace, checkpoint_every)
        return phi

    @staticmethod
    def backward(ctx, grad_out):
        (vel,) = ctx.saved_tensors
        win, dt, steps, interp, backtrace, every = ctx.cfg
        vel = vel.detach().contiguous()
        lam = grad_out.detach().contiguous().clone()
        rec = adjoint_record(vel, win, dt, interp, backtrace)
        hsum = torch.zeros(3 * win.nodes, dtype=vel.dtype, device=vel.device)
        other = torch.empty_like(lam)
        nseg = len(ctx.ckpts)
        for seg in reversed(range(nseg)):
            s0 = seg * every
            s1 = min(s0 + every, steps)
            states = [ctx.ckpts[seg]]
            for _ in range(s0, s1 - 1):
                states.append(advect_sl(states[-1], vel, win, dt, interp, backtrace))
            for s in reversed(range(s0, s1)):
                lam_prev = other.zero_()
                adjoint_step(states[s - s0], rec, lam, win, dt, lam_prev, hsum, interp)
                lam, other = lam_prev, lam
        gvel = adjoint_finish(vel, hsum, win, dt, torch.zeros_like(vel), backtrace)
        ctx.ckpts = None
        return lam, gvel, None, None, None, None, None, None


def advect_n(phi0, vel, win: GridWindow, dt: float, steps: int, interp="cubic", backtrace="rk2",
             checkpoint_every: int = 10):
    return AdvectN.apply(phi0, vel, win, dt, steps, interp, backtrace, checkpoint_every)


def heaviside(phi, eps: float):
    """H_eps with the reference's cosine delta as derivative (NormalDirection.c:44):
    H = (1 + phi/eps + sin(pi phi/eps)/pi)/2 on |phi| <= eps, 0 below, 1 above."""
    x = torch.clamp(phi / eps, -1.0, 1.0)
    return 0.5 * (1.0 + x + torch.sin(torch.pi * x) / torch.pi)
