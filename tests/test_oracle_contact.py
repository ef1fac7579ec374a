"""Pins the oracle's multi-body contact step (the FEMP contact algorithm,
SolveExplicitFEMP2DContact.jl:311-387, on the USL grid) with checks that do
not retype its formulas: before the solids meet it reduces to the
single-field step (grid bit for bit); the correction conserves momentum; a
head-on collision rebounds without the disks passing through each other
(which independent velocity fields without the correction would do); a
frictionless oblique impact spins the disks far less than a sticking one
(high friction, or the single shared field); a third solid at a node is
flagged; a rigid wall (its own normal, its velocity as the centre-of-mass
velocity, never corrected) changes nothing from afar, rebounds a disk
without letting it cross, keeps tangential momentum when frictionless and
pushes a resting disk when it moves.  In 3D (the MUSL step with
the normal-only correction) likewise: bitwise the single-field step while
apart, momentum conserved, a rebound without interpenetration where the
single shared field keeps the spheres together."""
import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.engine import make_material, make_params
from jump_amd.engine import JumpError
from oracle.pyoracle import OracleEngine


def _engine(scene):
    o = OracleEngine(scene.params)
    scene.load(o)
    return o


def _body_stats(o, s):
    a = o.get_solid(s)
    m = a["mass"]
    c = (a["pos"] * m[:, None]).sum(0) / m.sum()
    v = (a["vel"] * m[:, None]).sum(0) / m.sum()
    r = a["pos"] - c
    dv = a["vel"] - v
    w = float(((r[:, 0] * dv[:, 1] - r[:, 1] * dv[:, 0]) * m).sum() / ((r * r).sum(1) * m).sum())
    return c, v, w, a


def test_contact_grid_equals_single_field_before_the_solids_meet():
    kw = dict(gap_cells=6, steps=1)
    a = _engine(scenes.disks_contact(contact=True, **kw))
    b = _engine(scenes.disks_contact(contact=False, **kw))
    a.step(1, 0.0)
    b.step(1, 0.0)
    ga, gb = a.get_grid(), b.get_grid()
    for k in ("mass", "momentum0", "momentum", "force"):
        assert np.array_equal(ga[k], gb[k]), k
    a.step(149, 0.0)
    b.step(149, 0.0)
    for s in range(2):
        sa, sb = a.get_solid(s), b.get_solid(s)
        for k in ("pos", "vel", "F", "stress"):
            scale = np.max(np.abs(sb[k]))
            assert np.max(np.abs(sa[k] - sb[k])) <= 1e-12 * scale, (s, k)


@pytest.mark.parametrize("offset,friction", [(0.0, 0.0), (0.15, 0.0), (0.15, 0.3)])
def test_contact_conserves_momentum(offset, friction):
    o = _engine(scenes.disks_contact(offset=offset, friction=friction))
    ref = None
    for _ in range(8):
        o.step(100, 0.0)
        P = np.zeros(2)
        mv = 0.0
        for s in range(2):
            a = o.get_solid(s)
            P += (a["vel"] * a["mass"][:, None]).sum(0)
            mv += float((np.abs(a["vel"]).sum(1) * a["mass"]).sum())
        ref = mv if ref is None else ref
        assert np.max(np.abs(P)) <= 1e-12 * ref, P
    assert o.flags() == 0


def test_head_on_disks_rebound_without_passing_through():
    sc = scenes.disks_contact()
    o = _engine(sc)
    r = float(np.max(np.abs(sc.solids[0]["pos"][:, 0] - sc.solids[0]["pos"][:, 0].mean())))
    gaps = []
    for _ in range(12):
        o.step(100, 0.0)
        c0, v0, _, _ = _body_stats(o, 0)
        c1, v1, _, _ = _body_stats(o, 1)
        gaps.append(c1[0] - c0[0])
    assert min(gaps) > 0.9 * 2 * r
    assert v0[0] < -0.05 and v1[0] > 0.05  # rebounded
    assert abs(v0[1]) < 1e-12 and abs(v1[1]) < 1e-12  # symmetric impact stays on the line


def test_frictionless_oblique_impact_transmits_little_spin():
    spins = {}
    for name, kw in (("smooth", dict(friction=0.0)), ("rough", dict(friction=1e6)),
                     ("single", dict(contact=False))):
        o = _engine(scenes.disks_contact(offset=0.15, **kw))
        o.step(1000, 0.0)
        spins[name] = _body_stats(o, 0)[2]
    assert spins["rough"] > 0.3 and spins["single"] > 0.3, spins
    assert abs(spins["smooth"]) < 0.1 * spins["rough"], spins


def test_third_solid_at_a_node_is_flagged():
    sc = scenes.disks_contact(gap_cells=1)
    extra = dict(sc.solids[0])
    pos = sc.solids[0]["pos"].copy()
    pos[:, 1] += 0.5 * (sc.solids[1]["pos"][:, 0].mean() - pos[:, 0].mean())
    pos[:, 0] += 0.5 * (sc.solids[1]["pos"][:, 0].mean() - sc.solids[0]["pos"][:, 0].mean())
    extra["pos"] = pos
    extra["vel"] = np.zeros_like(pos)
    extra["vel"][:, 1] = -0.1
    sc.solids.append(extra)
    o = _engine(sc)
    o.step(300, 0.0)
    assert o.flags() & _ffi.JUMP_FLAG_CONTACT_OVERFLOW


def _wall_top(sc):
    return float(sc.solids[1]["pos"][:, 1].max())


def test_rigid_wall_far_away_changes_nothing():
    """While the disk is more than a stencil from the wall, no node holds
    both: the run equals the one without the wall, bit for bit."""
    a = _engine(scenes.disk_wall(gap_cells=8, steps=1))
    b = _engine(scenes.disk_wall(gap_cells=8, steps=1, wall=False))
    a.step(40, 0.0)
    b.step(40, 0.0)
    sa, sb = a.get_solid(0), b.get_solid(0)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    ga, gb = a.get_grid(), b.get_grid()
    for k in ("mass", "momentum0", "momentum", "force"):
        assert np.array_equal(ga[k], gb[k]), k  # the rigid solid adds no mass or momentum


@pytest.mark.parametrize("basis", [_ffi.BASIS_LINEAR, _ffi.BASIS_QUAD_BSPLINE])
def test_disk_rebounds_from_a_fixed_rigid_wall(basis):
    """Normal impact on a fixed wall: the disk rebounds (its vertical
    velocity changes sign) without its particles crossing the wall's
    surface, and the wall does not move."""
    sc = scenes.disk_wall(basis=basis, fixed=True)
    o = _engine(sc)
    top = _wall_top(sc)
    wall0 = sc.solids[1]["pos"].copy()
    lows = []
    for _ in range(10):
        o.step(100, 0.0)
        lows.append(float(o.get_solid(0)["pos"][:, 1].min()))
    c, v, _, _ = _body_stats(o, 0)
    assert v[1] > 0.05, v  # rebounded (it came in at -0.1)
    assert abs(v[0]) < 1e-12  # the symmetric impact stays on its line
    assert min(lows) > top, (min(lows), top)
    assert np.array_equal(o.get_solid(1)["pos"], wall0)


def test_frictionless_wall_keeps_tangential_momentum():
    """An oblique impact on a frictionless wall: the correction removes only
    the normal component at the contact nodes, so the disk's tangential
    momentum is what it would be without the wall (no body force, internal
    forces sum to zero); a sticking wall (friction 1e6) takes much of it."""
    res = {}
    for fr in (0.0, 1e6):
        sc = scenes.disk_wall(v=(0.08, -0.1), fixed=True, friction=fr)
        o = _engine(sc)
        m = sc.solids[0]["mass"]
        px0 = float((m * sc.solids[0]["vel"][:, 0]).sum())
        o.step(1000, 0.0)
        a = o.get_solid(0)
        res[fr] = float((a["mass"] * a["vel"][:, 0]).sum()) / px0
        assert float(a["vel"][:, 1].mean()) > 0.0  # rebounded
    assert abs(res[0.0] - 1.0) < 1e-12, res
    assert res[1e6] < 0.8, res


def test_sticking_wall_drags_the_disk_with_its_own_velocity():
    """With sticking friction the wall's tangential velocity is the contact
    nodes' centre-of-mass velocity: a wall sliding at 0.1 drags a falling
    disk along; the same wall marked fixed has zero velocity and drags
    nothing."""
    vx = {}
    for fixed in (False, True):
        o = _engine(scenes.disk_wall(wall_v=(0.1, 0.0), fixed=fixed, friction=1e6))
        o.step(1000, 0.0)
        vx[fixed] = _body_stats(o, 0)[1][0]
    assert vx[False] > 0.02, vx
    assert abs(vx[True]) < 1e-12, vx


def test_moving_rigid_platen_pushes_a_resting_disk():
    """A platen rising at 0.2 meets a resting disk: the disk is driven
    upwards and stays above the platen's moving surface."""
    sc = scenes.disk_wall(v=(0.0, 0.0), wall_v=(0.0, 0.2), gap_cells=1)
    o = _engine(sc)
    top0 = _wall_top(sc)
    t = 0.0
    for _ in range(6):
        t, _ = o.step(100, t)
        top = top0 + 0.2 * t
        assert float(o.get_solid(0)["pos"][:, 1].min()) > top - 0.25 / 32
    c, v, _, _ = _body_stats(o, 0)
    assert v[1] > 0.1, v
    assert abs(v[0]) < 1e-12


# ---------------------------------------------------------------- 3D (MUSL)
def _stats3(o, s):
    a = o.get_solid(s)
    m = a["mass"]
    return (a["pos"] * m[:, None]).sum(0) / m.sum(), (a["vel"] * m[:, None]).sum(0) / m.sum(), a


def test_contact3d_grid_equals_single_field_before_the_solids_meet():
    kw = dict(gap_cells=6, steps=1)
    a = _engine(scenes.spheres_contact(contact=True, **kw))
    b = _engine(scenes.spheres_contact(contact=False, **kw))
    a.step(1, 0.0)
    b.step(1, 0.0)
    ga, gb = a.get_grid(), b.get_grid()
    for k in ("mass", "momentum0", "momentum", "momentum2", "force"):
        assert np.array_equal(ga[k], gb[k]), k
    a.step(99, 0.0)
    b.step(99, 0.0)
    for s in range(2):
        sa, sb = a.get_solid(s), b.get_solid(s)
        for k in ("pos", "vel", "F", "stress"):
            assert np.array_equal(sa[k], sb[k]), (s, k)  # same sums in the same order while apart


@pytest.mark.parametrize("offset", [0.0, 0.15])
def test_contact3d_conserves_momentum_and_rebounds(offset):
    sc = scenes.spheres_contact(gap_cells=4, offset=offset)
    o = _engine(sc)
    ref = sum(float((np.abs(s["vel"]).sum(1) * s["mass"]).sum()) for s in sc.solids)
    overlaps = []
    for _ in range(12):
        o.step(100, 0.0)
        P = np.zeros(3)
        for s in range(2):
            a = o.get_solid(s)
            P += (a["vel"] * a["mass"][:, None]).sum(0)
        assert np.max(np.abs(P)) <= 1e-12 * ref, P
        c0, v0, a0 = _stats3(o, 0)
        c1, v1, a1 = _stats3(o, 1)
        overlaps.append(a0["pos"][:, 0].max() - a1["pos"][:, 0].min())
    if offset == 0.0:
        assert max(overlaps) < 0.0  # the spheres never interpenetrate
        assert v0[0] < -0.05 and v1[0] > 0.05  # and rebound
    assert o.flags() == 0


def test_contact3d_frictionless_oblique_impact_transmits_little_spin():
    """An oblique impact spins the spheres about z through the single shared
    field (it sticks); the normal-only correction transmits no tangential
    impulse."""
    def spin(o):
        a = o.get_solid(0)
        m = a["mass"]
        r = a["pos"] - (a["pos"] * m[:, None]).sum(0) / m.sum()
        dv = a["vel"] - (a["vel"] * m[:, None]).sum(0) / m.sum()
        return float(((r[:, 0] * dv[:, 1] - r[:, 1] * dv[:, 0]) * m).sum() / ((r[:, 0] ** 2 + r[:, 1] ** 2) * m).sum())
    w = {}
    for contact in (True, False):
        o = _engine(scenes.spheres_contact(gap_cells=4, offset=0.15, contact=contact))
        o.step(1200, 0.0)
        w[contact] = spin(o)
    assert w[False] > 0.3, w
    assert abs(w[True]) < 0.1 * w[False], w
