"""Reverse mode of the 3D MUSL step (jump_adjoint_*) on the GPU, for a
random linear functional L = sum pos_bar . x_T + vel_bar . v_T of the final
state: gradients with respect to E, nu, yield stress fy, hardening k1 and a
uniform change of the initial velocity against
  - the dual-number (ForwardDiff-style) run of the CPU restatement, whose
    forward is the GPU's bit for bit (Hooke, J2) or to 1e-10 (Neo-Hookean):
    relative 1e-9 of the gradient's scale;
  - central differences of the GPU's own forward step: relative 1e-6;
and bitwise independence of the checkpoint interval."""
import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.adjoint import rrule
from jump_amd.engine import Engine
from oracle.pyoracle import OracleEngine

pytestmark = pytest.mark.gpu

NSTEPS = 8
CASES = [(_ffi.BASIS_GIMP, "j2"), (_ffi.BASIS_LINEAR, "j2"), (_ffi.BASIS_GIMP, "neohookean"),
         (_ffi.BASIS_LINEAR, "elastic")]


def _scene(basis, material):
    return scenes.block3d(nodes=(9, 9, 9), basis=basis, material=material, steps=NSTEPS, seed=5)


def _bars(sc):
    rng = np.random.default_rng(11)
    n = sc.n_particles
    return rng.normal(size=(n, 3)), rng.normal(size=(n, 3))


def _gpu_gradient(sc, pb, vb, K):
    e = Engine(sc.params)
    sc.load(e)
    (pos, vel, t, rc), pullback = rrule(e, NSTEPS, checkpoint_every=K)
    assert rc == 0
    g = pullback(pb, vb)
    e.close()
    return g


def _loss(sc, pb, vb, **dp):
    m = sc.solids[0]["material"]
    keep = (m.E, m.nu, m.fy, m.k1)
    m.E += dp.get("E", 0.0)
    m.nu += dp.get("nu", 0.0)
    m.fy += dp.get("fy", 0.0)
    m.k1 += dp.get("k1", 0.0)
    v0 = sc.solids[0]["vel"].copy()
    if "v" in dp:
        sc.solids[0]["vel"] = v0 + dp["v"]
    e = Engine(sc.params)
    sc.load(e)
    e.step(NSTEPS, 0.0)
    a = e.get_solid(0)
    e.close()
    m.E, m.nu, m.fy, m.k1 = keep
    sc.solids[0]["vel"] = v0
    return float((pb * a["pos"]).sum() + (vb * a["vel"]).sum())


@pytest.mark.parametrize("basis,material", CASES)
def test_adjoint3d_matches_dual_numbers(hip_lib, basis, material):
    print(f"adjoint3d vs dual basis={basis} material={material}", flush=True)
    sc = _scene(basis, material)
    pb, vb = _bars(sc)
    g = _gpu_gradient(sc, pb, vb, 3)
    o = OracleEngine(sc.params)
    sc.load(o)
    want = o.dual_gradient(0, NSTEPS, pb, vb)
    o.close()
    got = {"E": g["E"][0], "nu": g["nu"][0], "fy": g["fy"][0], "k1": g["k1"][0]}
    v0 = want.pop("v0")
    for k in range(3):
        got[f"vel{k}"] = g["v0"][0, k]
        want[f"vel{k}"] = v0[k]
    print({k: (got[k], want[k]) for k in got})
    for k in got:
        assert abs(got[k] - want[k]) <= 1e-9 * max(abs(want[k]), 1e-300) + 1e-300, (k, got[k], want[k])
    if material == "j2":
        assert got["fy"] != 0.0 and got["k1"] != 0.0  # plastic flow took place


def test_adjoint3d_matches_central_differences(hip_lib):
    sc = _scene(_ffi.BASIS_GIMP, "j2")
    pb, vb = _bars(sc)
    g = _gpu_gradient(sc, pb, vb, 0)
    m = sc.solids[0]["material"]
    checks = [("E", 1e-6 * m.E, g["E"][0], 1e-6), ("nu", 1e-6, g["nu"][0], 1e-6),
              ("fy", 1e-6 * m.fy, g["fy"][0], 1e-6), ("k1", 1e-6 * m.k1, g["k1"][0], 1e-4)]
    for key, h, got, rtol in checks:
        fd = (_loss(sc, pb, vb, **{key: h}) - _loss(sc, pb, vb, **{key: -h})) / (2 * h)
        print(key, got, fd)
        assert abs(got - fd) <= rtol * abs(fd), (key, got, fd)
    for k in range(3):
        dv = np.zeros(3)
        dv[k] = 1e-6
        fd = (_loss(sc, pb, vb, v=dv) - _loss(sc, pb, vb, v=-dv)) / 2e-6
        assert abs(g["v0"][0, k] - fd) <= 1e-6 * abs(fd), (k, g["v0"][0, k], fd)


@pytest.mark.parametrize("ppc", [1, 4])
def test_adjoint3d_lanes_per_node(hip_lib, ppc):
    """The backward sweep's node gathers at 1 and 4 lanes per node (1 and 4
    particles per cell and direction; the blocks above run 2): against the
    dual-number restatement at relative 1e-9, and independent of the
    checkpoint interval."""
    print(f"adjoint3d lanes ppc={ppc}", flush=True)
    sc = scenes.block3d(nodes=(9, 9, 9), basis=_ffi.BASIS_GIMP, material="j2", steps=NSTEPS, seed=6, ppc=ppc)
    pb, vb = _bars(sc)
    g = _gpu_gradient(sc, pb, vb, 3)
    g2 = _gpu_gradient(sc, pb, vb, NSTEPS)
    for k in ("E", "nu", "fy", "k1", "v0"):
        assert np.array_equal(g[k], g2[k]), k
    o = OracleEngine(sc.params)
    sc.load(o)
    want = o.dual_gradient(0, NSTEPS, pb, vb)
    o.close()
    got = {"E": g["E"][0], "nu": g["nu"][0], "fy": g["fy"][0], "k1": g["k1"][0]}
    v0 = want.pop("v0")
    for k in range(3):
        got[f"vel{k}"] = g["v0"][0, k]
        want[f"vel{k}"] = v0[k]
    print({k: (got[k], want[k]) for k in got})
    for k in got:
        assert abs(got[k] - want[k]) <= 1e-9 * max(abs(want[k]), 1e-300) + 1e-300, (k, got[k], want[k])


def test_adjoint3d_checkpoint_interval_does_not_change_the_gradient(hip_lib):
    sc = _scene(_ffi.BASIS_GIMP, "j2")
    pb, vb = _bars(sc)
    res = [_gpu_gradient(sc, pb, vb, K) for K in (1, 3, NSTEPS, 0)]
    for r in res[1:]:
        for k in ("E", "nu", "fy", "k1", "v0"):
            assert np.array_equal(r[k], res[0][k]), k


def test_taylor_bar_gradients_reduced_size(hip_lib):
    """The Taylor bar at reduced size (particle spacing x4: 7,700 particles,
    11 x 41 x 11 nodes, 60 steps, a snapshot every 20): the inverse problem's
    loss L = sum |x_T - x*|^2 against a run with other (E, fy, hardening);
    its gradient by the reverse mode equals the dual-number run of the CPU
    restatement to 1e-9 and central differences of the GPU forward to 1e-6,
    and plastic flow takes part (dL/dfy != 0)."""
    print("taylor bar reduced size", flush=True)
    steps = 60
    tgt = scenes.taylor_bar(scale=4.0, steps=steps, E=125e9, fy=450e6, H=150e6)
    e = Engine(tgt.params)
    tgt.load(e)
    e.step(steps, 0.0)
    xstar = e.get_solid(0)["pos"]
    e.close()
    sc = scenes.taylor_bar(scale=4.0, steps=steps)
    e = Engine(sc.params)
    sc.load(e)
    (pos, vel, t, rc), pullback = rrule(e, steps, checkpoint_every=20)
    assert rc == 0
    pb = 2.0 * (pos - xstar)
    g = pullback(pb)
    e.close()
    o = OracleEngine(sc.params)
    sc.load(o)
    want = o.dual_gradient(0, steps, pb)
    o.close()
    got = {"E": g["E"][0], "nu": g["nu"][0], "fy": g["fy"][0], "k1": g["k1"][0], "vy": g["v0"][0, 1]}
    ref = {"E": want["E"], "nu": want["nu"], "fy": want["fy"], "k1": want["k1"], "vy": want["v0"][1]}
    print({k: (got[k], ref[k]) for k in got})
    for k in got:
        assert abs(got[k] - ref[k]) <= 1e-9 * abs(ref[k]), (k, got[k], ref[k])
    assert got["fy"] != 0.0

    def loss(**dp):
        m = sc.solids[0]["material"]
        keep = (m.E, m.fy, m.k1)
        m.E += dp.get("E", 0.0)
        m.fy += dp.get("fy", 0.0)
        m.k1 += dp.get("k1", 0.0)
        e = Engine(sc.params)
        sc.load(e)
        e.step(steps, 0.0)
        x = e.get_solid(0)["pos"]
        e.close()
        m.E, m.fy, m.k1 = keep
        return float(((x - xstar) ** 2).sum())

    m = sc.solids[0]["material"]
    for key, h, rtol in (("E", 1e-6 * m.E, 1e-6), ("fy", 1e-6 * m.fy, 1e-6), ("k1", 1e-5 * m.k1, 1e-4)):
        fd = (loss(**{key: h}) - loss(**{key: -h})) / (2 * h)
        print(key, got[key], fd)
        assert abs(got[key] - fd) <= rtol * abs(fd), (key, got[key], fd)


@pytest.mark.parametrize("basis,material", [(_ffi.BASIS_GIMP, "j2"), (_ffi.BASIS_LINEAR, "neohookean")])
def test_recorded_steps_equal_plain_steps_3d(hip_lib, basis, material):
    """The 3D recorded run and the backward sweep's rebuilds step from record
    to record: G2P-1 writes the new velocity into the next record, the
    re-projection reads it there, and G2P-2 writes the rest, with alpha and
    the plastic strain after every step.  At every checkpoint interval
    (segments of one step, uneven last segments, the whole run) the recorded
    run ends in the plain run's state bitwise, and two sweeps give the same
    gradient, the second rebuilding every segment."""
    sc = _scene(basis, material)
    e = Engine(sc.params)
    sc.load(e)
    e.step(NSTEPS, 0.0)
    ref = e.get_solid(0)
    e.close()
    pb, vb = _bars(sc)
    res = []
    for K in (1, 3, NSTEPS):
        e = Engine(sc.params)
        sc.load(e)
        (pos, vel, t, rc), pullback = rrule(e, NSTEPS, checkpoint_every=K)
        assert rc == 0
        assert np.array_equal(pos, ref["pos"]) and np.array_equal(vel, ref["vel"]), K
        g1 = pullback(pb, vb)
        g2 = pullback(pb, vb)
        for k in ("E", "nu", "v0", "fy", "k1"):
            assert np.array_equal(g1[k], g2[k]), (K, k)
        res.append(g1)
        e.close()
    for r in res[1:]:
        for k in ("E", "nu", "v0", "fy", "k1"):
            assert np.array_equal(r[k], res[0][k]), k
