"""GPU parity of the 2D MUSL step with the blended mass matrix
(SolveExplicitSolidMechanics2D.jl:25-329) against the oracle.

The oracle solves B v = p with a dense LU as the reference does; the device
solves the same system by the Chebyshev iteration with its error bounded by
1e-14 of the solution, so the comparison is to a relative 1e-10 of each
field's scale (not bitwise).  Beyond the oracle's 4096-node limit the checks are the
size-independent closed forms: a uniformly translating stress-free body
keeps its velocity and moves by dt v, and a uniform body force adds dt g."""
import os
import sys

import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.engine import Engine, JumpError, make_material, make_params

pytestmark = pytest.mark.gpu

RTOL = 1e-10
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _close(a, b, what, rtol=RTOL, atol=0.0):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    err = float(np.max(np.abs(a - b))) if b.size else 0.0
    assert err <= rtol * scale + atol + 1e-300, f"{what}: max err {err:.3e} vs scale {scale:.3e}"


def _compare(g, o, nsolids, atol=None):
    """Every field to RTOL of its scale; `atol` adds absolute floors for
    fields that are roundoff in the run (a rigidly translating body's strain)."""
    atol = atol or {}
    for s in range(nsolids):
        a, b = g.get_solid(s), o.get_solid(s)
        for k in ("pos", "vel", "volume", "F", "strain", "stress", "mass"):
            _close(a[k], b[k], f"solid {s} {k}", atol=atol.get(k, 0.0))
        sa, sb = g.get_state(s), o.get_state(s)
        for k in ("alpha", "alphadot", "vm", "pstrain"):
            _close(sa[k], sb[k], f"solid {s} {k}")
    ga, gb = g.get_grid(), o.get_grid()
    for k in ("mass", "momentum0", "momentum", "force"):
        _close(ga[k], gb[k], f"grid {k}", atol=atol.get("grid " + k, 0.0))
    _close(g.energies(), o.energies(), "energies")


def _run(scene, steps):
    from oracle.pyoracle import OracleEngine
    g = Engine(scene.params)
    scene.load(g)
    o = OracleEngine(scene.params)
    scene.load(o)
    tg, rg = g.step(steps, 0.0)
    to, ro = o.step(steps, 0.0)
    assert tg == to and rg == ro == 0
    assert g.flags() == o.flags() == 0
    return g, o


@pytest.mark.parametrize("material", ["elastic", "neohookean", "j2"])
@pytest.mark.parametrize("basis", [_ffi.BASIS_LINEAR, _ffi.BASIS_QUAD_BSPLINE, _ffi.BASIS_CUBIC_BSPLINE,
                                   _ffi.BASIS_GIMP])
def test_musl2d_block(hip_lib, basis, material):
    """Six steps: with the linear basis the reference's algorithm turns a
    particle's det F negative on this block from the eighth step on."""
    _say(f"musl2d block basis={basis} material={material}")
    sc = scenes.block2d(cells=16, basis=basis, material=material, steps=6, algorithm=_ffi.ALG_MUSL, alpha=0.99)
    g, o = _run(sc, sc.steps)
    _compare(g, o, 1)


@pytest.mark.parametrize("alpha", [1.0, 0.0])
def test_musl2d_block_flip_fraction(hip_lib, alpha):
    _say(f"musl2d block alpha={alpha}")
    sc = scenes.block2d(cells=16, basis=_ffi.BASIS_QUAD_BSPLINE, steps=12, algorithm=_ffi.ALG_MUSL, alpha=alpha)
    g, o = _run(sc, sc.steps)
    _compare(g, o, 1)


def test_musl2d_binned_order(hip_lib):
    _say("musl2d block, binned order")
    from oracle.pyoracle import OracleEngine
    sc = scenes.block2d(cells=16, basis=_ffi.BASIS_QUAD_BSPLINE, steps=12, algorithm=_ffi.ALG_MUSL, alpha=0.99)
    sc.params.order = _ffi.ORDER_BINNED
    g = Engine(sc.params)
    sc.load(g)
    sc.params.order = _ffi.ORDER_REFERENCE
    o = OracleEngine(sc.params)
    sc.load(o)
    g.step(sc.steps, 0.0)
    o.step(sc.steps, 0.0)
    _compare(g, o, 1)


def test_musl2d_cpdi_two_disks_from_gmsh(hip_lib):
    """Main_TwoDisksCPDI.jl as the example runs it: CPDI-Q4 disks from
    disk.msh on the 11 x 11 grid, MUSL(1.0); 300 steps.  The domain corners
    stay where they were built (the corner update reads a nodal momentum the
    step never fills, SolveExplicitSolidMechanics2D.jl:289-309), so the
    disks translate without deforming: strain, stress and the nodal forces
    are roundoff and are compared to an absolute 1e-12 (stress: E times
    that, forces: 1e-12 N)."""
    _say("musl2d cpdi two disks from disk.msh")
    from jump_amd import algorithm, basis, bodyforce, grid, material, problem, solid
    from oracle.pyoracle import OracleEngine
    path = os.path.join(GOLD, "disk.msh")
    g = grid.Grid2D(0.0, 1.5, 0.0, 1.5, 11, 11)
    m = material.ElasticMaterial(1000.0, 0.3, 1000.0, 0, 0, 1000)
    s1, s2 = solid.Solid2D(path, m), solid.Solid2D(path, m)
    solid.assign_velocity(s1, (0.1, 0.1))
    solid.assign_velocity(s2, (-0.1, -0.1))
    solid.move_cpdi(s1, (0.45, 0.45))
    solid.move_cpdi(s2, (0.45 + 0.6, 0.45 + 0.6))
    data = {"total_time": 3.0, "dt": 1e-3, "time": 0.0, "bodyforce": bodyforce.ConstantBodyForce2D((0.0, 0.0))}
    engines = [problem.build_engine(g, [s1, s2], basis.CPDIQ4Basis(), algorithm.MUSL(1.0), data, engine=cls)
               for cls in (Engine, OracleEngine)]
    for e in engines:
        _, rc = e.step(300, 0.0)
        assert rc == 0
    assert engines[0].flags() == 0
    _compare(engines[0], engines[1], 2, atol={"strain": 1e-12, "stress": 1e-9, "grid force": 1e-12})


def test_musl2d_two_rings_cpdi_from_gmsh(hip_lib):
    """Main_TwoRings_CPDI.jl as the example runs it: two rings from the
    reference's ring.msh (Neo-Hookean, mm-s units) at +-30 m/s on the
    200 x 120 mm grid of 81 x 41 nodes, CPDI-Q4, MUSL(0.99), dt = 0.2 dx/c;
    20 steps (the oracle's dense LU takes 0.2 s a step here).  As in the disks
    case the domains stay put, so strain, stress and forces are roundoff."""
    _say("musl2d cpdi two rings from ring.msh")
    import math
    from jump_amd import algorithm, basis, bodyforce, grid, material, problem, solid
    from oracle.pyoracle import OracleEngine
    density, K, mu = 1010e-12, 121.7, 26.1
    E = 9 * K * mu / (3 * K + mu)
    nu = E / 2 / mu - 1.0
    w, l, vel = 120.0, 200.0, 30e3
    g = grid.Grid2D(0, l, 0, w, 81, 41)
    m1, m2 = material.NeoHookeanMaterial(E, nu, density, 10000), material.NeoHookeanMaterial(E, nu, density, 10000)
    c = math.sqrt((m1.lambda_ + 2 * m1.mu) / m1.density)
    path = os.path.join(GOLD, "ring.msh")
    s1, s2 = solid.Solid2D(path, m1), solid.Solid2D(path, m2)
    solid.assign_velocity(s1, (vel, 0.0))
    solid.assign_velocity(s2, (-vel, 0.0))
    solid.move_cpdi(s1, (l / 4, w / 2))
    solid.move_cpdi(s2, (0.75 * l, 0.5 * w))
    data = {"total_time": 0.0047, "dt": 0.2 * g.dx / c, "time": 0.0,
            "bodyforce": bodyforce.ConstantBodyForce2D((0.0, 0.0))}
    engines = [problem.build_engine(g, [s1, s2], basis.CPDIQ4Basis(), algorithm.MUSL(0.99), data, engine=cls)
               for cls in (Engine, OracleEngine)]
    for e in engines:
        _, rc = e.step(20, 0.0)
        assert rc == 0
    assert engines[0].flags() == 0
    _compare(engines[0], engines[1], 2, atol={"strain": 1e-12, "stress": 1e-10, "grid force": 1e-10})


def test_musl2d_cpdi_tubes_stretching(hip_lib):
    """CPDI-Q4 tube walls (jittered corners), no platens, on the 49 x 49
    grid, MUSL(0.99), 40 steps: an affine stretching velocity field (x out,
    y in) plus seeded noise.  The time step is a tenth of the tubes' usual
    0.2 cell crossing: the blended consistent-mass step is only stable well
    below that (at 0.05 cell crossings per step the block diverges)."""
    _say("musl2d cpdi tubes, stretching")
    sc = scenes.cpdi_tubes(rings=(2, 2), cells=48, nr=3, nt=48, steps=40, platens=False)
    pos = sc.solids[0]["pos"]
    c = pos.mean(0)
    rng = np.random.default_rng(7)
    sc.solids[0]["vel"] = 2.0 * np.stack([(pos[:, 0] - c[0]) / 0.01, -(pos[:, 1] - c[1]) / 0.01], 1) + \
        rng.normal(scale=0.1, size=pos.shape)
    sc.params.algorithm = _ffi.ALG_MUSL
    sc.params.alpha = 0.99
    sc.params.dt *= 0.1
    g, o = _run(sc, sc.steps)
    assert np.max(np.abs(o.get_solid(0)["strain"])) > 1e-6
    _compare(g, o, 1)


@pytest.mark.parametrize("cells", [300, 600])
@pytest.mark.parametrize("basis", [_ffi.BASIS_LINEAR, _ffi.BASIS_QUAD_BSPLINE, _ffi.BASIS_CUBIC_BSPLINE])
def test_musl2d_large_grid_closed_forms(hip_lib, basis, cells):
    """A 301 x 301 grid (90601 nodes, beyond the dense restatement) and a
    601 x 601 grid (1444 node tiles, so tiles of every kind: interior,
    right and top edges): a stress-free block translating under gravity
    gains exactly dt g per step at every particle, whatever the FLIP
    fraction."""
    _say(f"musl2d large grid basis={basis} cells={cells}")
    dx = 1.0 / cells
    f = dx / 2
    xs = np.arange(0.2 + f / 2, 0.8, f)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel()], 1)
    n = pos.shape[0]
    vol = np.full(n, f * f)
    v = np.array([0.3, -0.2])
    gv = np.array([0.0, -9.81])
    dt = 1e-5
    p = make_params(2, basis, _ffi.ALG_MUSL, dt, (0.0, 0.0), (1.0, 1.0), (cells + 1, cells + 1), alpha=0.7,
                    body=(gv[0], gv[1], 0.0))
    e = Engine(p)
    mat = make_material(_ffi.MAT_ELASTIC, E=1.0e6, nu=0.3, density=1000.0)
    e.add_solid(mat, pos, np.tile(v, (n, 1)), 1000.0 * vol, vol, vol)
    steps = 3
    e.step(steps, 0.0)
    assert e.flags() == 0
    a = e.get_solid(0)
    vt = v + steps * dt * gv
    xt = pos + sum(dt * (v + k * dt * gv) for k in range(1, steps + 1))
    assert np.max(np.abs(a["vel"] - vt)) <= 1e-12
    assert np.max(np.abs(a["pos"] - xt)) <= 1e-12
    e.close()


def test_musl2d_refusals(hip_lib):
    _say("musl2d refusals")
    from jump_amd import adjoint
    from jump_amd.engine import Batch
    p = make_params(2, _ffi.BASIS_LINEAR, _ffi.ALG_MUSL, 1e-4, (0.0, 0.0), (1.0, 1.0), (11, 11), contact=True)
    with pytest.raises(JumpError):
        Engine(p)
    sc = scenes.block2d(cells=8, steps=1, algorithm=_ffi.ALG_MUSL)
    g = Engine(sc.params)
    rig = make_material(_ffi.MAT_RIGID, density=1000.0)
    with pytest.raises(JumpError):
        g.add_solid(rig, np.array([[0.5, 0.5]]), np.zeros((1, 2)), np.ones(1), np.ones(1) * 1e-3)
    sc.load(g)
    with pytest.raises(JumpError):
        Batch([g])
    with pytest.raises(JumpError):
        adjoint.rrule(g, 2)
    g.close()
