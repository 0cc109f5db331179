"""Sensitivities of the conjugate heat transfer case (hom3g_cht_*).

The forward map is the reference's coupled step (coupling::apply, then
cns.solve(), then heat.solve(); coupling.hpp:47-63).  The objective is the
time-integrated wall heat flux out of the solid through the conjugate
interface, J = sum_n dt_heat^n Q^{n+1}; the parameters are the solid's
per-cell volumetric source.  Tangent and adjoint run on a frozen dt
schedule (the CFL-limited schedule of a reference run), so J is a smooth
function of the source.

With ``obs`` (one value per internal solid cell) the objective is instead
the final-state misfit J = 1/2 sum (T_N - obs)^2 of an inverse-source
problem (hom3g_cht_misfit_*).

``WallHeatFlux`` and ``SourceMisfit`` wrap the adjoint as
``torch.autograd.Function``s (torch is only imported when they are used).
"""
import ctypes as C

import numpy as np

from ._lib import check, lib

_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class Snapshot:
    """Host copy of both solvers' state and clocks, to rerun from the same start."""

    def __init__(self, *solvers):
        self.items = [(s, s.lhs.copy(), s.time(), s.dt(), s.step()) for s in solvers]

    def restore(self):
        for s, q, t, dt, st in self.items:
            s.set_lhs(q)
            s.reset_time(t, dt, st)


def _obs(heat, obs):
    o = np.ascontiguousarray(obs, np.float64)
    if o.size != heat.first_gc:
        raise ValueError(f"obs needs one value per internal solid cell ({heat.first_gc}), got {o.size}")
    return o


def cht_objective(cns, heat, Tr0, nsteps, dts=None, obs=None):
    """Run nsteps coupled steps; returns (J, dts used)."""
    if obs is not None:
        if dts is None:
            raise ValueError("the misfit objective needs a dt schedule")
        d = np.ascontiguousarray(dts, np.float64)
        o = _obs(heat, obs)
        J = C.c_double()
        check(lib().hom3g_cht_misfit_objective(cns.ctx.h, cns.id, heat.id, float(Tr0), int(nsteps), _ptr(d),
                                               _ptr(o), C.byref(J)))
        return J.value, d
    J = C.c_double()
    out = np.empty(nsteps)
    d = None if dts is None else np.ascontiguousarray(dts, np.float64)
    check(lib().hom3g_cht_objective(cns.ctx.h, cns.id, heat.id, float(Tr0), int(nsteps), _ptr(d),
                                    _ptr(out) if dts is None else None, C.byref(J)))
    return J.value, (out if dts is None else d)


def cht_tangent(cns, heat, Tr0, dts, dsource, obs=None):
    """FP64 forward-mode dJ along dsource (per internal solid cell)."""
    d = np.ascontiguousarray(dts, np.float64)
    v = np.ascontiguousarray(dsource, np.float64)
    J, dJ = C.c_double(), C.c_double()
    if obs is not None:
        o = _obs(heat, obs)
        check(lib().hom3g_cht_misfit_tangent(cns.ctx.h, cns.id, heat.id, float(Tr0), d.size, _ptr(d), _ptr(o),
                                             _ptr(v), C.byref(J), C.byref(dJ)))
    else:
        check(lib().hom3g_cht_tangent(cns.ctx.h, cns.id, heat.id, float(Tr0), d.size, _ptr(d), _ptr(v), C.byref(J),
                                      C.byref(dJ)))
    return J.value, dJ.value


_CKPT_MEMORY = {"auto": 0, "device": 1, "host": 2}


def cht_adjoint(cns, heat, Tr0, dts, ckpt_every=20, obs=None, ckpt_memory="auto"):
    """Discrete adjoint: (J, dJ/dsource, ms forward sweep, ms reverse sweep).
    ckpt_every > 0 checkpoints every that many steps and recomputes each
    segment once; 0 keeps every step in device memory and recomputes
    nothing; -S uses binomial checkpointing (revolve's schedule) with at
    most S saved states.  ckpt_memory: where the checkpoints of
    ckpt_every > 0 live, "device", "host" (pinned host memory) or "auto"
    (device memory while it fits)."""
    check(lib().hom3g_set_checkpoint_memory(cns.ctx.h, _CKPT_MEMORY[ckpt_memory]))
    d = np.ascontiguousarray(dts, np.float64)
    g = np.empty(heat.first_gc)
    J = C.c_double()
    mf, mr = C.c_float(), C.c_float()
    if obs is not None:
        o = _obs(heat, obs)
        check(lib().hom3g_cht_misfit_adjoint(cns.ctx.h, cns.id, heat.id, float(Tr0), d.size, _ptr(d),
                                             int(ckpt_every), _ptr(o), C.byref(J), _ptr(g), C.byref(mf),
                                             C.byref(mr)))
    else:
        check(lib().hom3g_cht_adjoint(cns.ctx.h, cns.id, heat.id, float(Tr0), d.size, _ptr(d), int(ckpt_every),
                                      C.byref(J), _ptr(g), C.byref(mf), C.byref(mr)))
    return J.value, g, mf.value, mr.value


def WallHeatFlux():
    """torch.autograd.Function: source (internal solid cells) -> J.

    forward(source, cns, heat, Tr0, dts, snapshot, ckpt_every) restores the
    snapshot, sets the source, runs the adjoint sweep (forward with
    checkpoints, then reverse) and keeps dJ/dsource for backward."""
    import torch

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, source, cns, heat, Tr0, dts, snapshot, ckpt_every=20):
            snapshot.restore()
            heat.set_source(source.detach().cpu().double().numpy())
            J, g, _, _ = cht_adjoint(cns, heat, Tr0, dts, ckpt_every)
            ctx.save_for_backward(torch.from_numpy(g).to(source.dtype))
            return source.new_tensor(J)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            return gJ * g.to(gJ.device), None, None, None, None, None, None

    return _Fn


def SourceMisfit():
    """torch.autograd.Function: source (internal solid cells) -> the
    final-state misfit 1/2 sum (T_N - obs)^2 of an inverse-source problem.

    forward(source, cns, heat, Tr0, dts, obs, snapshot, ckpt_every) restores
    the snapshot, sets the source, runs the adjoint sweep and keeps
    dJ/dsource for backward."""
    import torch

    class _Fn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, source, cns, heat, Tr0, dts, obs, snapshot, ckpt_every=20):
            snapshot.restore()
            heat.set_source(source.detach().cpu().double().numpy())
            J, g, _, _ = cht_adjoint(cns, heat, Tr0, dts, ckpt_every, obs=obs)
            ctx.save_for_backward(torch.from_numpy(g).to(source.dtype))
            return source.new_tensor(J)

        @staticmethod
        def backward(ctx, gJ):
            (g,) = ctx.saved_tensors
            return gJ * g.to(gJ.device), None, None, None, None, None, None, None

    return _Fn
