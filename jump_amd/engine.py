"""Thin object wrapper over the C ABI of include/jump_mpm.h.

`Engine` owns one jump_ctx of libjumphip.so.  Arrays cross the boundary in the
reference's AoS layout: positions (n, d), matrices (n, d*d) column-major per
particle (the memory of Vector{SMatrix{d,d,Float64}}).
"""
import ctypes
import os

import numpy as np

from . import _ffi


class JumpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"jump error {code}: {msg}")
        self.code = code


def make_params(dim, basis, algorithm, dt, lo, hi, nodes, tolerance=0.0, alpha=1.0, body=(0.0, 0.0, 0.0),
                ghostcell=False, order=_ffi.ORDER_REFERENCE, contact=False, friction=0.0):
    p = _ffi.JumpParams()
    p.dim = dim
    p.basis = basis
    p.algorithm = algorithm
    p.ghostcell = 1 if ghostcell else 0
    p.tolerance = tolerance
    p.alpha = alpha
    p.dt = dt
    p.xmin, p.ymin = lo[0], lo[1]
    p.xmax, p.ymax = hi[0], hi[1]
    p.zmin = lo[2] if dim == 3 else 0.0
    p.zmax = hi[2] if dim == 3 else 0.0
    p.nx, p.ny = nodes[0], nodes[1]
    p.nz = nodes[2] if dim == 3 else 1
    for k in range(3):
        p.body[k] = body[k] if k < len(body) else 0.0
    p.order = order
    p.contact = 1 if contact else 0
    p.friction = friction
    return p


def make_material(kind, E=0.0, nu=0.0, density=0.0, fy=0.0, k1=0.0, A=0.0, B=0.0, C=0.0, n=1.0, epsdot0=1.0,
                  m=1.0, chi=0.0, Cp=1.0, cellsize=1.0, vx=0.0, vy=0.0, fixed=False):
    mt = _ffi.JumpMaterial()
    mt.kind = kind
    mt.fixed = 1 if fixed else 0
    mt.E, mt.nu, mt.density = E, nu, density
    mt.fy, mt.k1 = fy, k1
    mt.A, mt.B, mt.C, mt.n, mt.epsdot0, mt.m, mt.chi, mt.Cp, mt.cellsize = A, B, C, n, epsdot0, m, chi, Cp, cellsize
    mt.vx, mt.vy = vx, vy
    return mt


def _c(a, dtype=np.float64):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


class Engine:
    """One simulation context.  `lib`/`prefix` select the implementation:
    the HIP library by default."""

    def __init__(self, params, device=0, lib=None, prefix="jump_", create=None):
        self.lib = lib if lib is not None else _ffi.load()
        self.prefix = prefix
        self.dim = params.dim
        self.params = params
        self.nodes = (params.nx, params.ny, params.nz if params.dim == 3 else 1)
        self.nn = int(np.prod(self.nodes))
        self.counts = []
        h = ctypes.c_void_p()
        if create is None:
            rc = self.lib.jump_ctx_create(ctypes.byref(params), device, ctypes.byref(h))
        else:
            rc = create(ctypes.byref(params), ctypes.byref(h))
        if rc != 0:
            raise JumpError(rc, "context creation failed")
        self.h = h

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc):
        if rc != 0:
            msg = self._f("last_error")(self.h)
            raise JumpError(rc, msg.decode() if msg else "")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self._f("ctx_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_solid(self, material, pos, vel, mass, volume, volume0=None, F=None, strain=None, stress=None,
                  corners=None, elems=None, lp=None):
        d = self.dim
        pos = _c(pos).reshape(-1, d)
        n = pos.shape[0]
        keep = dict(
            pos=pos, vel=_c(vel).reshape(n, d), mass=_c(mass).reshape(n), volume=_c(volume).reshape(n),
            volume0=_c(volume if volume0 is None else volume0).reshape(n),
            F=None if F is None else _c(F).reshape(n, d * d),
            strain=None if strain is None else _c(strain).reshape(n, d * d),
            stress=None if stress is None else _c(stress).reshape(n, d * d),
            corners=None if corners is None else _c(corners).reshape(-1, 2),
            elems=None if elems is None else _c(elems, np.int64).reshape(n, 4))
        ds = _ffi.JumpSolidDesc()
        ds.n = n
        for f in ("pos", "vel", "mass", "volume", "volume0", "F", "strain", "stress"):
            setattr(ds, f, _ffi.dptr(keep[f]))
        ds.ncorners = 0 if keep["corners"] is None else keep["corners"].shape[0]
        ds.corners = _ffi.dptr(keep["corners"])
        ds.elems = _ffi.i64ptr(keep["elems"])
        if lp is not None:
            for k in range(d):
                ds.lp[k] = float(lp[k]) if np.ndim(lp) else float(lp)
        idx = ctypes.c_int32()
        self._check(self._f("add_solid")(self.h, ctypes.byref(ds), ctypes.byref(material), ctypes.byref(idx)))
        self.counts.append(n)
        return idx.value

    def set_fixed(self, mask):
        """mask: (nn, d) nonzero = fixed (the reference's fixedNodes')."""
        m = np.ascontiguousarray(mask, dtype=np.uint8).reshape(self.nn, self.dim)
        self._check(self._f("set_fixed")(self.h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))

    def step(self, nsteps, t=0.0):
        """Runs nsteps steps from time t; returns (t_new, rc) where rc is 0 or
        JUMP_ERR_NUMERIC when particle error flags were raised."""
        tt = ctypes.c_double(t)
        rc = self._f("step")(self.h, int(nsteps), ctypes.byref(tt))
        if rc not in (0, _ffi.JUMP_ERR_NUMERIC):
            self._check(rc)
        return tt.value, rc

    def flags(self):
        f = ctypes.c_uint32()
        self._check(self._f("flags")(self.h, ctypes.byref(f)))
        return f.value

    def get_solid(self, s):
        d, n = self.dim, self.counts[s]
        out = dict(pos=np.empty((n, d)), vel=np.empty((n, d)), mass=np.empty(n), volume=np.empty(n),
                   F=np.empty((n, d * d)), strain=np.empty((n, d * d)), stress=np.empty((n, d * d)))
        self._check(self._f("get_solid")(self.h, s, *[_ffi.dptr(out[k]) for k in
                                                       ("pos", "vel", "mass", "volume", "F", "strain", "stress")]))
        return out

    def get_state(self, s):
        d, n = self.dim, self.counts[s]
        out = dict(alpha=np.empty(n), alphadot=np.empty(n), vm=np.empty(n),
                   pstrain=np.empty((n, 3 if d == 2 else 9)))
        self._check(self._f("get_solid_state")(self.h, s, *[_ffi.dptr(out[k]) for k in
                                                             ("alpha", "alphadot", "vm", "pstrain")]))
        return out

    def set_body_force(self, kind, prm):
        """jump_set_body_force: _ffi.BODY_CONSTANT with the dim components, or
        _ffi.BODY_VORTEX2D with (G, T, Ri, Ro, rho0, mu, xcenter, ycenter)."""
        q = np.ascontiguousarray(prm, dtype=np.float64).ravel()
        self._check(self._f("set_body_force")(self.h, int(kind), _ffi.dptr(q), int(q.size)))

    def get_affine(self, s):
        """APIC: the solid's affine matrices C_p, (n, 4) column-major."""
        out = np.empty((self.counts[s], 4))
        self._check(self._f("get_affine")(self.h, s, _ffi.dptr(out)))
        return out

    def set_affine(self, s, cmat):
        c = np.ascontiguousarray(cmat, dtype=np.float64)
        if c.shape != (self.counts[s], 4):
            raise ValueError(f"cmat must have shape ({self.counts[s]}, 4), got {c.shape}")
        self._check(self._f("set_affine")(self.h, s, _ffi.dptr(c)))

    def get_grid(self):
        d, nn = self.dim, self.nn
        out = dict(mass=np.empty(nn), momentum0=np.empty((nn, d)), momentum=np.empty((nn, d)),
                   momentum2=np.empty((nn, d)), force=np.empty((nn, d)))
        self._check(self._f("get_grid")(self.h, *[_ffi.dptr(out[k]) for k in
                                                  ("mass", "momentum0", "momentum", "momentum2", "force")]))
        return out

    def energies(self):
        ke, se = ctypes.c_double(), ctypes.c_double()
        self._check(self._f("energies")(self.h, ctypes.byref(ke), ctypes.byref(se)))
        return ke.value, se.value

    def dump_lammps(self, path, outs=()):
        """Queue a LAMMPS particle dump of the current state (OvitoOutput,
        Output.jl:146-260) appended to `path`; formatted and written by the
        library's writer thread (HIP library only)."""
        cols = outs if isinstance(outs, str) else " ".join(outs)
        self._check(self.lib.jump_dump_lammps(self.h, os.fsencode(path), cols.encode()))

    def dump_wait(self):
        """Block until every queued dump is on disk."""
        self._check(self.lib.jump_dump_wait(self.h))

    def profile(self, enable=True):
        """Enable (and reset) per-phase HIP-event timing (HIP library only)."""
        self._check(self.lib.jump_profile(self.h, 1 if enable else 0))

    def profile_read(self):
        out = {}
        for i, name in enumerate(_ffi.PHASES):
            ms, cnt = ctypes.c_double(), ctypes.c_int64()
            self._check(self.lib.jump_profile_read(self.h, i, ctypes.byref(ms), ctypes.byref(cnt)))
            out[name] = (ms.value, cnt.value)
        return out

    def checkpoint(self):
        """Snapshot of the particle state (bytes); see restore()."""
        n = ctypes.c_int64()
        self._check(self.lib.jump_checkpoint_bytes(self.h, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        self._check(self.lib.jump_checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def restore(self, blob):
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        self._check(self.lib.jump_checkpoint_load(self.h, buf, len(blob)))

    def bins(self):
        """Diagnostic: per-particle stencil origin (HIP library only)."""
        n = sum(self.counts)
        keys = np.empty(n, dtype=np.int32)
        self._check(self.lib.jump_get_bins(self.h, keys.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return keys


class Batch:
    """Many 2D USL Engines stepped together (jump_batch_*).  The engines keep
    their particle counts; after step() call sync() before reading them."""

    def __init__(self, engines, device=0):
        self.lib = engines[0].lib
        self.engines = list(engines)
        arr = (ctypes.c_void_p * len(engines))(*[e.h.value for e in engines])
        h = ctypes.c_void_p()
        rc = self.lib.jump_batch_create(arr, len(engines), device, ctypes.byref(h))
        if rc != 0:
            raise JumpError(rc, "batch creation failed")
        self.h = h
        self.t = np.zeros(len(engines))

    def _check(self, rc):
        if rc != 0:
            msg = self.lib.jump_batch_last_error(self.h)
            raise JumpError(rc, msg.decode() if msg else "")
        return rc

    def step(self, nsteps):
        t = np.ascontiguousarray(self.t)
        rc = self.lib.jump_batch_step(self.h, int(nsteps), _ffi.dptr(t))
        self.t = t
        if rc not in (0, _ffi.JUMP_ERR_NUMERIC):
            self._check(rc)
        return rc

    def sync(self):
        self._check(self.lib.jump_batch_sync(self.h))

    def options(self, graph=True, autotune=True, shape=0):
        """HIP-graph replay of the step and the gather-shape autotuner (both
        on by default; neither changes results); without the autotuner every
        group uses `shape` (_ffi.P2G_SHAPE_NAMES)."""
        self._check(self.lib.jump_batch_options(self.h, 1 if graph else 0, 1 if autotune else 0, int(shape)))

    def tuning(self):
        """The autotuner's record: per group, (chosen shape, ms per launch of
        each shape)."""
        ng = ctypes.c_int32()
        self.lib.jump_batch_tuning(self.h, 0, ctypes.byref(ng), None, None)
        out = []
        for g in range(ng.value):
            sh = ctypes.c_int32()
            ms = np.zeros(_ffi.P2G_SHAPES)
            self._check(self.lib.jump_batch_tuning(self.h, g, None, ctypes.byref(sh), _ffi.dptr(ms)))
            out.append((sh.value, ms))
        return out

    def energies(self):
        ke, se = np.zeros(len(self.engines)), np.zeros(len(self.engines))
        self._check(self.lib.jump_batch_energies(self.h, _ffi.dptr(ke), _ffi.dptr(se)))
        return ke, se

    def close(self):
        if getattr(self, "h", None):
            self.lib.jump_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
