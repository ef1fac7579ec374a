"""Reverse mode of the 2D USL and 3D MUSL steps through the C ABI
(jump_adjoint_*).

`rrule(engine, nsteps, checkpoint_every)` is the shape of a ChainRulesCore
rrule around solve_explicit_dynamics_2D/3D (SolveExplicitSolidMechanics2D.jl:
335, SolveExplicitSolidMechanics3D.jl:14): it runs the steps and returns the
final particle positions and velocities with a pullback that maps their
adjoints to gradients with respect to each solid's E, nu, initial velocity
and, for J2 solids in 3D, yield stress fy and hardening k1.
checkpoint_every = 0 lets the library pick the interval from the free device
memory (the whole run when it fits, so the backward sweep recomputes
nothing); `pullback.checkpoint_every` is the interval used.
"""
import ctypes

import numpy as np

from . import _ffi
from .engine import JumpError


def rrule(engine, nsteps, checkpoint_every=0, t=0.0):
    lib = engine.lib
    tt = ctypes.c_double(t)
    rc = lib.jump_adjoint_forward(engine.h, int(nsteps), int(checkpoint_every), ctypes.byref(tt))
    if rc not in (0, _ffi.JUMP_ERR_NUMERIC):
        engine._check(rc)
    pos = np.concatenate([engine.get_solid(s)["pos"] for s in range(len(engine.counts))])
    vel = np.concatenate([engine.get_solid(s)["vel"] for s in range(len(engine.counts))])

    def pullback(pos_bar, vel_bar=None):
        ns = len(engine.counts)
        npart = int(sum(engine.counts))
        d = getattr(engine, "dim", 2)
        # the library reads dim doubles per particle from each array
        pb = np.ascontiguousarray(pos_bar, dtype=np.float64)
        if pb.shape != (npart, d):
            raise ValueError(f"pos_bar must have shape ({npart}, {d}), got {pb.shape}")
        vb = None
        if vel_bar is not None:
            vb = np.ascontiguousarray(vel_bar, dtype=np.float64)
            if vb.shape != (npart, d):
                raise ValueError(f"vel_bar must have shape ({npart}, {d}), got {vb.shape}")
        dE, dnu, dv0 = np.zeros(ns), np.zeros(ns), np.zeros((ns, d))
        rc2 = lib.jump_adjoint_backward(engine.h, _ffi.dptr(pb), _ffi.dptr(vb), _ffi.dptr(dE), _ffi.dptr(dnu),
                                        _ffi.dptr(dv0))
        if rc2 != 0:
            engine._check(rc2)
        dfy, dk1 = np.zeros(ns), np.zeros(ns)
        engine._check(lib.jump_adjoint_plastic(engine.h, _ffi.dptr(dfy), _ffi.dptr(dk1)))
        return {"E": dE, "nu": dnu, "v0": dv0, "fy": dfy, "k1": dk1}

    k = ctypes.c_int32()
    engine._check(lib.jump_adjoint_interval(engine.h, ctypes.byref(k)))
    pullback.checkpoint_every = k.value
    return (pos, vel, tt.value, rc), pullback
