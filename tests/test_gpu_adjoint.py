"""Reverse mode of the 2D USL step (jump_adjoint_*): gradients of
L = sum |x_T - x*|^2 with respect to E, nu and each solid's initial velocity
against central finite differences of the forward step on the GPU (the
forward step is itself bitwise equal to the CPU restatement)."""
import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.adjoint import rrule
from jump_amd.engine import Engine

pytestmark = pytest.mark.gpu

NSTEPS = 20


def _scene(basis, material):
    return scenes.block2d(cells=16, basis=basis, material=material, steps=NSTEPS, seed=3)


def _loss(basis, material, target, dE=0.0, dnu=0.0, dv=(0.0, 0.0)):
    sc = _scene(basis, material)
    m = sc.solids[0]["material"]
    m.E += dE
    m.nu += dnu
    sc.solids[0]["vel"] = sc.solids[0]["vel"] + np.asarray(dv)
    e = Engine(sc.params)
    sc.load(e)
    e.step(NSTEPS, 0.0)
    pos = e.get_solid(0)["pos"]
    e.close()
    return float(np.sum((pos - target) ** 2))


@pytest.mark.parametrize("basis,material", [(_ffi.BASIS_LINEAR, "elastic"), (_ffi.BASIS_LINEAR, "neohookean"),
                                            (_ffi.BASIS_GIMP, "elastic"), (_ffi.BASIS_GIMP, "neohookean")])
def test_adjoint_gradients_match_central_differences(hip_lib, basis, material):
    sc = _scene(basis, material)
    target = sc.solids[0]["pos"] + np.array([0.004, -0.003])
    e = Engine(sc.params)
    sc.load(e)
    (pos, vel, t, rc), pullback = rrule(e, NSTEPS, checkpoint_every=6)
    assert rc == 0
    g = pullback(2.0 * (pos - target))
    E, nu = sc.solids[0]["material"].E, sc.solids[0]["material"].nu
    fd = {}
    h = 1e-5 * E
    fd["E"] = (_loss(basis, material, target, dE=h) - _loss(basis, material, target, dE=-h)) / (2 * h)
    h = 1e-5  # the central difference in nu is most accurate here (truncation above, rounding below)
    fd["nu"] = (_loss(basis, material, target, dnu=h) - _loss(basis, material, target, dnu=-h)) / (2 * h)
    for k in range(2):
        dv = np.zeros(2)
        dv[k] = 1e-6
        fd[f"v{k}"] = (_loss(basis, material, target, dv=dv) - _loss(basis, material, target, dv=-dv)) / 2e-6
    got = {"E": g["E"][0], "nu": g["nu"][0], "v0": g["v0"][0, 0], "v1": g["v0"][0, 1]}
    print(basis, material, {k: (got[k], fd[k]) for k in fd})
    for k in fd:
        assert abs(got[k] - fd[k]) <= 1e-6 * abs(fd[k]) + 1e-12, (k, got[k], fd[k])


def test_checkpoint_interval_does_not_change_the_gradient(hip_lib):
    sc = _scene(_ffi.BASIS_LINEAR, "neohookean")
    target = sc.solids[0]["pos"]
    res = []
    for K in (1, 7, NSTEPS, 0):
        e = Engine(sc.params)
        sc.load(e)
        (pos, vel, t, rc), pullback = rrule(e, NSTEPS, checkpoint_every=K)
        assert pullback.checkpoint_every == (K or NSTEPS)  # automatic: the small run fits whole
        res.append(pullback(2.0 * (pos - target), vel_bar=np.ones_like(vel)))
        after = e.get_solid(0)["pos"]
        assert np.array_equal(after, pos)  # the context keeps the recorded final state
    for r in res[1:]:
        for k in ("E", "nu", "v0"):
            print(k, r[k], res[0][k])
            assert np.array_equal(r[k], res[0][k]), k  # every adjoint sum has a fixed order


def test_stepping_after_the_recorded_run_discards_it(hip_lib):
    """The pullback belongs to the state the recorded run ended in: after a
    further jump_step it is refused (JUMP_ERR_STATE) and the particles keep
    the stepped state instead of being rewound."""
    from jump_amd.engine import JumpError
    sc = _scene(_ffi.BASIS_LINEAR, "elastic")
    e = Engine(sc.params)
    sc.load(e)
    (pos, vel, t, rc), pullback = rrule(e, 5, checkpoint_every=2)
    assert rc == 0
    e.step(3, t)
    moved = e.get_solid(0)["pos"]
    assert not np.array_equal(moved, pos)
    with pytest.raises(JumpError) as ei:
        pullback(np.ones_like(pos))
    assert ei.value.code == _ffi.JUMP_ERR_STATE
    assert np.array_equal(e.get_solid(0)["pos"], moved)
    e.close()


@pytest.mark.parametrize("basis,material", [(_ffi.BASIS_LINEAR, "neohookean"), (_ffi.BASIS_GIMP, "elastic")])
def test_adjoint2d_matches_dual_numbers(hip_lib, basis, material):
    """The 2D reverse mode against the dual-number run of the CPU
    restatement (whose forward equals the GPU's): relative 1e-9."""
    from oracle.pyoracle import OracleEngine
    sc = _scene(basis, material)
    rng = np.random.default_rng(12)
    n = sc.n_particles
    pb, vb = rng.normal(size=(n, 2)), rng.normal(size=(n, 2))
    e = Engine(sc.params)
    sc.load(e)
    (pos, vel, t, rc), pullback = rrule(e, NSTEPS, checkpoint_every=7)
    g = pullback(pb, vb)
    e.close()
    o = OracleEngine(sc.params)
    sc.load(o)
    want = o.dual_gradient(0, NSTEPS, pb, vb)
    got = {"E": g["E"][0], "nu": g["nu"][0], "v0": g["v0"][0, 0], "v1": g["v0"][0, 1]}
    ref = {"E": want["E"], "nu": want["nu"], "v0": want["v0"][0], "v1": want["v0"][1]}
    print({k: (got[k], ref[k]) for k in got})
    for k in got:
        assert abs(got[k] - ref[k]) <= 1e-9 * abs(ref[k]), (k, got[k], ref[k])


@pytest.mark.parametrize("ppc", [1, 4])
def test_adjoint2d_lanes_per_node(hip_lib, ppc):
    """The node gathers of the backward sweep split each node over as many
    lanes as the run has particles per cell and direction (1, 2 or 4, fixed
    when the run is recorded); the 2 of the blocks above, and here 1 and 4,
    each against the dual-number restatement at relative 1e-9, and the
    gradient independent of the checkpoint interval."""
    from oracle.pyoracle import OracleEngine
    sc = scenes.block2d(cells=12, basis=_ffi.BASIS_GIMP, material="neohookean", steps=NSTEPS, seed=7, ppc=ppc)
    rng = np.random.default_rng(5)
    n = sc.n_particles
    pb, vb = rng.normal(size=(n, 2)), rng.normal(size=(n, 2))
    res = []
    for K in (7, NSTEPS):
        e = Engine(sc.params)
        sc.load(e)
        (pos, vel, t, rc), pullback = rrule(e, NSTEPS, checkpoint_every=K)
        assert rc == 0
        res.append(pullback(pb, vb))
        e.close()
    for k in ("E", "nu", "v0"):
        assert np.array_equal(res[0][k], res[1][k]), k
    o = OracleEngine(sc.params)
    sc.load(o)
    want = o.dual_gradient(0, NSTEPS, pb, vb)
    g = res[0]
    got = {"E": g["E"][0], "nu": g["nu"][0], "v0": g["v0"][0, 0], "v1": g["v0"][0, 1]}
    ref = {"E": want["E"], "nu": want["nu"], "v0": want["v0"][0], "v1": want["v0"][1]}
    print(ppc, {k: (got[k], ref[k]) for k in got})
    for k in got:
        assert abs(got[k] - ref[k]) <= 1e-9 * abs(ref[k]), (k, got[k], ref[k])


@pytest.mark.parametrize("basis", [_ffi.BASIS_LINEAR, _ffi.BASIS_GIMP])
def test_recorded_steps_equal_plain_steps(hip_lib, basis):
    """The 2D recorded run and the backward sweep's rebuilds step from record
    to record (each step reads its state from one record and writes the
    next).  Two colliding rings (the reverse mode refuses rigid solids): at
    every checkpoint interval (segments of one step, uneven last segments,
    the whole run) the recorded run ends in the plain run's state bitwise,
    and the gradient is the same for every interval and for a second sweep,
    which rebuilds every segment."""
    n = 30
    sc = scenes.rings2d(cells=40, basis=basis, platen=False, steps=n)
    e = Engine(sc.params)
    sc.load(e)
    e.step(n, 0.0)
    ns = len(sc.solids)
    ref_pos = np.concatenate([e.get_solid(s)["pos"] for s in range(ns)])
    ref_vel = np.concatenate([e.get_solid(s)["vel"] for s in range(ns)])
    e.close()
    res = []
    for K in (1, 4, 7, n):
        e = Engine(sc.params)
        sc.load(e)
        (pos, vel, t, rc), pullback = rrule(e, n, checkpoint_every=K)
        assert np.array_equal(pos, ref_pos) and np.array_equal(vel, ref_vel), K
        g1 = pullback(np.ones_like(pos), vel_bar=np.ones_like(vel))
        g2 = pullback(np.ones_like(pos), vel_bar=np.ones_like(vel))
        for k in ("E", "nu", "v0"):
            assert np.all(np.isfinite(g1[k])), (K, k)
            assert np.array_equal(g1[k], g2[k]), (K, k)
        res.append(g1)
        e.close()
    for r in res[1:]:
        for k in ("E", "nu", "v0"):
            assert np.array_equal(r[k], res[0][k]), k
