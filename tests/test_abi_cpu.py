"""CPU checks of the boundary: the HIP library loads (no GPU call) and
exports every function include/jump_mpm.h declares; the host mirror of the
reference API (ranges, builders, Dirichlet masks) matches the reference."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from jump_amd import (Grid2D, Grid3D, Solid2D, buildParticleForBlock, buildParticleForCircle,
                      buildParticleForSphere, ElasticMaterial, fix_Dirichlet_grid, julia_range, _ffi)
from jump_amd.fix import julia_float

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _declared():
    src = open(os.path.join(ROOT, "include", "jump_mpm.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(jump_\w+)\s*\(", src, re.M)))


def test_header_declares_the_boundary():
    names = _declared()
    for must in ("jump_ctx_create", "jump_ctx_destroy", "jump_add_solid", "jump_set_fixed", "jump_step",
                 "jump_get_solid", "jump_get_solid_state", "jump_get_grid", "jump_energies", "jump_last_error",
                 "jump_flags", "jump_checkpoint_bytes", "jump_checkpoint_save", "jump_checkpoint_load",
                 "jump_batch_create", "jump_batch_step", "jump_batch_sync", "jump_batch_energies",
                 "jump_batch_destroy"):
        assert must in names


def test_library_loads_and_exports_every_declared_symbol():
    lib = _ffi.load()  # raises if the HIP library is not built
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.jump_abi_version() == 1
    assert sorted(_ffi.EXPORTED) == _declared()


def test_oracle_exports_the_same_entry_points(oracle_lib):
    for name in _declared():
        if name in ("jump_get_bins", "jump_abi_version", "jump_profile", "jump_profile_read",
                    "jump_checkpoint_bytes", "jump_checkpoint_save", "jump_checkpoint_load",
                    "jump_energies_noalloc") or \
                name.startswith(("jump_batch_", "jump_adjoint_", "jump_dump_")):
            continue  # diagnostics, device-state transfer, batching, reverse mode, dumps: GPU library only
        assert hasattr(oracle_lib, name.replace("jump_", "orc_")), name


def test_julia_ranges_match_fixture():
    d = json.load(open(os.path.join(GOLD, "julia_ranges.json")))
    for v in d.values():
        r = julia_range(v["start"], v["step"], v["stop"])
        assert len(r) == v["length"]
        assert r[0] == v["first"] and r[-1] == v["last"]
    assert len(julia_range(0.0, 0.1, 1.0)) == 11
    assert len(julia_range(0.0125, 0.025, 1.0 - 0.0125)) == 40


def test_builders_match_reference_counts():
    grid = Grid2D(0.0, 1.0, 0.0, 1.0, 21, 21)
    f = grid.dx / 2
    c1 = buildParticleForCircle([0.2, 0.2], 0.2, f)
    assert c1.shape == (208, 2)
    # x fastest, as the reference's nested loops (Solid.jl:327-333)
    assert c1[1, 1] == c1[0, 1] and c1[1, 0] > c1[0, 0]
    b = buildParticleForBlock([0.0, 0.0, 0.0], 100e-3, 100e-3, 20e-3, 0.36e-3)
    assert b.shape[0] == 277 * 277 * 55
    s = buildParticleForSphere([0.0, 0.0, 0.0], 5e-3, 0.36e-3)
    assert s.shape[0] == 9328
    sol = Solid2D(c1, ElasticMaterial(1000.0, 0.3, 1000.0, 0, 0, len(c1)))
    assert sol.deformationGradient.shape == (208, 2, 2)
    sol.deformationGradient[3, 0, 1] = 2.5  # [a, b] view over column-major storage
    assert sol._F[3, 2] == 2.5


def test_dirichlet_faces():
    g = Grid2D(0.0, 1.0, 0.0, 1.0, 5, 4)
    fix_Dirichlet_grid(g, {"dirichlet_grid": [("bottom", (0, 1, 0)), ("left", (1, 0, 0))]})
    assert g.fixedNodes[1, 1:5].all() and not g.fixedNodes[0, 1:5].any()
    assert g.fixedNodes[0, [0, 5, 10, 15]].all()
    # `.= fix` assigns every component: the later "left" entry resets the
    # corner's y fix (Dirichlet.jl:58,70)
    assert list(g.fixedNodes[:, 0]) == [1, 0]
    with pytest.raises(ValueError):
        fix_Dirichlet_grid(g, {"dirichlet_grid": [("front", (1, 1, 1))]})
    g3 = Grid3D(0, 1, 0, 1, 0, 1, 3, 4, 5)
    fix_Dirichlet_grid(g3, {"dirichlet_grid": [("back", (1, 1, 1))]})
    assert g3.fixedNodes[:, -12:].all() and not g3.fixedNodes[:, :-12].any()


def test_energy_record_format():
    assert julia_float(1e-5) == "1.0e-5"
    assert julia_float(2.6) == "2.6"
    assert julia_float(0.0) == "0.0"


def test_grid_node_cap_is_refused_before_any_device_call():
    """Node-indexed int arrays are read with 32-bit byte offsets, so a grid
    of 2^30 - 1 nodes or more is refused (JUMP_ERR_UNSUPPORTED) by argument
    checks that run before the library touches a device."""
    from jump_amd.engine import make_params
    lib = _ffi.load()
    out = ctypes.c_void_p()
    for nodes in ((32768, 32768), (1024, 1024, 1024), (65535, 16385)):
        d = len(nodes)
        p = make_params(d, _ffi.BASIS_LINEAR if d == 2 else _ffi.BASIS_QUAD_BSPLINE,
                        _ffi.ALG_USL if d == 2 else _ffi.ALG_MUSL, 1e-3, (0.0,) * d, (1.0,) * d, nodes)
        assert lib.jump_ctx_create(ctypes.byref(p), 0, ctypes.byref(out)) == _ffi.JUMP_ERR_UNSUPPORTED, nodes
        assert not out.value
    p = make_params(2, _ffi.BASIS_LINEAR, _ffi.ALG_USL, 1e-3, (0.0, 0.0), (1.0, 1.0), (65536, 4))
    assert lib.jump_ctx_create(ctypes.byref(p), 0, ctypes.byref(out)) == _ffi.JUMP_ERR_UNSUPPORTED


class _StubLib:
    """Stands in for libjumphip.so: records the pullback call."""

    def __init__(self):
        self.backward_calls = 0

    def jump_adjoint_forward(self, h, n, k, t):
        return 0

    def jump_adjoint_interval(self, h, k):
        return 0

    def jump_adjoint_backward(self, h, pb, vb, dE, dnu, dv0):
        self.backward_calls += 1
        return 0

    def jump_adjoint_plastic(self, h, dfy, dk1):
        return 0


class _StubEngine:
    def __init__(self, counts):
        self.lib = _StubLib()
        self.h = None
        self.counts = list(counts)

    def get_solid(self, s):
        return {"pos": np.zeros((self.counts[s], 2)), "vel": np.zeros((self.counts[s], 2))}

    def _check(self, rc):
        assert rc == 0


def test_pullback_checks_adjoint_shapes():
    """The library reads 2 doubles per particle from pos_bar and vel_bar, so
    the pullback refuses arrays of any other shape before calling it."""
    from jump_amd.adjoint import rrule
    eng = _StubEngine([3, 2])
    _, pullback = rrule(eng, 4)
    for bad in (np.zeros((4, 2)), np.zeros(10), np.zeros((5, 3)), np.zeros((2, 5))):
        with pytest.raises(ValueError):
            pullback(bad)
        with pytest.raises(ValueError):
            pullback(np.zeros((5, 2)), bad)
    assert eng.lib.backward_calls == 0
    pullback(np.zeros((5, 2)), np.ones((5, 2)))
    pullback(np.zeros((5, 2)))
    assert eng.lib.backward_calls == 2
