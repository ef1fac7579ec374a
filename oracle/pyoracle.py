"""Python binding of the CPU restatement (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  It drives the oracle through the
same Engine wrapper and call sequence as the HIP library (entry points
prefixed `orc_` instead of `jump_`).
"""
import ctypes
import os
import subprocess

import numpy as np

from jump_amd import _ffi
from jump_amd.engine import Engine

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    _ffi._declare(lib, "orc_")
    lib.orc_ctx_create.restype = ctypes.c_int
    lib.orc_ctx_create.argtypes = [ctypes.POINTER(_ffi.JumpParams), ctypes.POINTER(ctypes.c_void_p)]
    lib.orc_bspline.restype = ctypes.c_int
    lib.orc_bspline.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]
    lib.orc_gimp1d.restype = ctypes.c_int
    lib.orc_gimp1d.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double)]
    lib.orc_support.restype = ctypes.c_int
    lib.orc_support.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_double)]
    lib.orc_stress.restype = ctypes.c_int
    lib.orc_stress.argtypes = [ctypes.c_int, ctypes.POINTER(_ffi.JumpMaterial)] + [ctypes.POINTER(ctypes.c_double)] * 3 + \
        [ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.orc_dual_gradient.restype = ctypes.c_int
    lib.orc_dual_gradient.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.orc_step_parallel.restype = ctypes.c_int
    lib.orc_step_parallel.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    lib.orc_det.restype = ctypes.c_double
    lib.orc_det.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    _LIB = lib
    return lib


class OracleEngine(Engine):
    def __init__(self, params):
        lib = load()
        super().__init__(params, lib=lib, prefix="orc_", create=lib.orc_ctx_create)

    def step_parallel(self, nsteps, t=0.0, threads=1):
        """3D MUSL steps on `threads` OpenMP threads (the many-core CPU
        baseline; node sums formed block by block)."""
        tt = ctypes.c_double(t)
        rc = self.lib.orc_step_parallel(self.h, int(nsteps), ctypes.byref(tt), int(threads))
        if rc not in (0, _ffi.JUMP_ERR_NUMERIC):
            self._check(rc)
        return tt.value, rc

    def dual_gradient(self, solid, nsteps, pos_bar, vel_bar=None):
        """dL/d(E, nu, fy, k1, v0) of solid `solid` for L = sum pos_bar . x +
        vel_bar . v after nsteps steps from the current state, by the
        restatement run on dual numbers (the context is not stepped)."""
        d = self.dim
        n = int(sum(self.counts))
        pb = np.ascontiguousarray(pos_bar, dtype=np.float64).reshape(n, d)
        vb = None if vel_bar is None else np.ascontiguousarray(vel_bar, dtype=np.float64).reshape(n, d)
        g = np.zeros(4 + d)
        rc = self.lib.orc_dual_gradient(self.h, int(solid), int(nsteps), _ffi.dptr(pb), _ffi.dptr(vb), _ffi.dptr(g))
        if rc not in (0, _ffi.JUMP_ERR_NUMERIC):
            self._check(rc)
        return {"E": g[0], "nu": g[1], "fy": g[2], "k1": g[3], "v0": g[4:].copy()}

    def support(self, s, p, grad=True):
        nd = self.dim
        near = np.zeros(64, dtype=np.int64)
        funcs = np.zeros(64)
        ders = np.zeros(64 * 3)
        cnt = self.lib.orc_support(self.h, s, p, 1 if grad else 0, near.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                   _ffi.dptr(funcs), _ffi.dptr(ders))
        if cnt < 0:
            return None
        return near[:cnt].copy(), funcs[:cnt].copy(), ders[:nd * cnt].reshape(cnt, nd).copy()


def bspline(family, btype, r):
    lib = load()
    N, dN = ctypes.c_double(), ctypes.c_double()
    rc = lib.orc_bspline(family, btype, r, ctypes.byref(N), ctypes.byref(dN))
    assert rc == 0
    return N.value, dN.value


def gimp1d(xi, h, lp):
    lib = load()
    N, dN = ctypes.c_double(), ctypes.c_double()
    lib.orc_gimp1d(xi, h, lp, ctypes.byref(N), ctypes.byref(dN))
    return N.value, dN.value


def stress(dim, material, eps, deps=None, F=None, J=1.0, dt=0.0, sig=None, state=None):
    """One constitutive update on full column-major tensors; returns
    (sigma, state) with state = [alpha, alphadot, vm, pstrain...]."""
    lib = load()
    dd = dim * dim
    eps = np.ascontiguousarray(eps, dtype=np.float64).reshape(dd)
    deps = np.zeros(dd) if deps is None else np.ascontiguousarray(deps, dtype=np.float64).reshape(dd)
    F = np.eye(dim).reshape(dd) if F is None else np.ascontiguousarray(F, dtype=np.float64).reshape(dd)
    sig = np.zeros(dd) if sig is None else np.array(sig, dtype=np.float64).reshape(dd)
    state = np.zeros(3 + (3 if dim == 2 else 9)) if state is None else np.array(state, dtype=np.float64)
    rc = lib.orc_stress(dim, ctypes.byref(material), _ffi.dptr(eps), _ffi.dptr(deps), _ffi.dptr(F), J, dt,
                        _ffi.dptr(sig), _ffi.dptr(state))
    assert rc == 0, rc
    return sig, state


def det(dim, A):
    lib = load()
    A = np.ascontiguousarray(A, dtype=np.float64).reshape(dim * dim)
    return lib.orc_det(dim, _ffi.dptr(A))
