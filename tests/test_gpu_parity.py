"""GPU parity: the HIP step (through the C ABI) against the CPU restatement
on the same seeded inputs, for every basis x material combination the
reference defines on this path, plus the Two Disks energy history fixture and
size-independent properties at the benchmark size.

Both sides evaluate every term in the reference's operation order (no FMA
contraction) and the GPU's P2G folds each node's contributions in the
reference's particle order, so particle and grid state are bit-identical for
materials built from +,-,*,/,sqrt (Hooke, J2).  Neo-Hookean (log) and
Johnson-Cook (pow) call the device math library, which may differ from the
host's in the last ulp: those agree within 1e-10 relative (normwise per field:
max|gpu - cpu| <= 1e-10 * max|cpu|).  Energies are reduced in a different
order on the GPU: relative 1e-12, or 2 n 2^-53 for n particles where that is
larger (the error bound of the restatement's recursive sum).  Binning is
bit-exact."""
import math
import os
import sys

import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.engine import Engine, JumpError

pytestmark = pytest.mark.gpu

RTOL = 1e-10
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _close(a, b, rtol=RTOL, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    scale = float(np.max(np.abs(b))) if b.size else 0.0
    err = float(np.max(np.abs(a - b))) if b.size else 0.0
    assert err <= rtol * scale + 1e-300, f"{what}: max err {err:.3e} vs scale {scale:.3e}"


def _run(scene, steps, hip_lib):
    from oracle.pyoracle import OracleEngine
    g = Engine(scene.params)
    scene.load(g)
    o = OracleEngine(scene.params)
    scene.load(o)
    tg, rg = g.step(steps, 0.0)
    to, ro = o.step(steps, 0.0)
    assert tg == to
    assert rg == ro
    assert g.flags() == o.flags()
    return g, o


def _same(a, b, exact, what):
    if exact:
        a, b = np.asarray(a), np.asarray(b)
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        assert not bad.any(), f"{what}: {int(bad.sum())} values differ, max {np.max(np.abs(a - b)[bad]):.3e}"
    else:
        _close(a, b, what=what)


def _compare(g, o, scene, grid=True, exact=True):
    for s in range(len(scene.solids)):
        a, b = g.get_solid(s), o.get_solid(s)
        for k in ("pos", "vel", "volume", "F", "strain", "stress", "mass"):
            _same(a[k], b[k], exact, f"solid {s} {k}")
        sa, sb = g.get_state(s), o.get_state(s)
        for k in ("alpha", "alphadot", "vm", "pstrain"):
            _same(sa[k], sb[k], exact, f"solid {s} {k}")
    if grid:
        ga, gb = g.get_grid(), o.get_grid()
        keys = ("mass", "momentum0", "momentum", "force") + (("momentum2",) if g.dim == 3 else ())
        for k in keys:
            _same(ga[k], gb[k], exact, f"grid {k}")
    ea, eb = g.energies(), o.energies()
    # the restatement sums particle by particle, the GPU in a two-level tree:
    # recursive summation of n terms is exact to (n - 1) u sum|x| (u = 2^-53)
    n = sum(len(s["mass"]) for s in scene.solids)
    _close(ea, eb, rtol=max(1e-12, 2.0 * n * 2.0 ** -53), what="energies")


def _exact(material):
    return material in ("elastic", "j2")


BASES_2D = [_ffi.BASIS_LINEAR, _ffi.BASIS_QUAD_BSPLINE, _ffi.BASIS_CUBIC_BSPLINE, _ffi.BASIS_GIMP]
BASES_3D = [_ffi.BASIS_LINEAR, _ffi.BASIS_QUAD_BSPLINE, _ffi.BASIS_CUBIC_BSPLINE, _ffi.BASIS_GIMP]


@pytest.mark.parametrize("material", ["elastic", "neohookean", "j2"])
@pytest.mark.parametrize("basis", BASES_2D)
def test_usl2d_block(hip_lib, basis, material):
    _say(f"usl2d block basis={basis} material={material}")
    sc = scenes.block2d(cells=16, basis=basis, material=material, steps=12)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=_exact(material))


@pytest.mark.parametrize("material", ["elastic", "neohookean", "j2", "jc", "jc_rate"])
@pytest.mark.parametrize("basis", BASES_3D)
def test_musl3d_block(hip_lib, basis, material):
    _say(f"musl3d block basis={basis} material={material}")
    sc = scenes.block3d(nodes=(11, 11, 11), basis=basis, material=material, steps=6)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=_exact(material))
    if material == "jc_rate":
        # the rate path ran: the filtered rate exceeds the reference rate
        # somewhere, so (1 + ratio)^C scaled the yield stress
        st = g.get_state(0)
        assert np.max(st["alphadot"]) > 1.0e-5 and np.max(st["alpha"]) > 0.0


@pytest.mark.parametrize("basis", [_ffi.BASIS_LINEAR, _ffi.BASIS_QUAD_BSPLINE, _ffi.BASIS_GIMP])
def test_usl2d_rings_rigid_platen(hip_lib, basis):
    _say(f"rings2d basis={basis}")
    sc = scenes.rings2d(cells=40, basis=basis, steps=60)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=False)
    sc = scenes.rings2d(cells=40, basis=basis, material="elastic", steps=60)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=True)


def test_usl2d_cpdi_tubes(hip_lib):
    _say("cpdi tubes")
    sc = scenes.cpdi_tubes(rings=(2, 2), cells=48, nr=3, nt=48, steps=15)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=False)


def test_musl3d_impact_small(hip_lib):
    _say("impact3d small")
    sc = scenes.impact3d(nodes=(36, 36, 22), plate_mm=(16.0, 16.0, 5.0), sphere_r_mm=2.0, steps=25)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=True)


def test_two_disks_energy_history(hip_lib):
    """Two disks (Main_TwoDisks.jl at 20x20 cells): 3500 steps, energies every
    10 steps against the committed record and the live restatement."""
    _say("two disks 3500 steps")
    from oracle.pyoracle import OracleEngine
    rec = np.load(os.path.join(GOLD, "two_disks_energy.npz"))["rec"]
    sc = scenes.two_disks()
    g = Engine(sc.params)
    sc.load(g)
    out = []
    t, _ = g.step(1, 0.0)
    out.append((0.0,) + g.energies())
    for c in range(1, 3500):
        tc = t
        t, _ = g.step(1, t)
        if c % 10 == 0:
            out.append((tc,) + g.energies())
    out = np.array(out)
    assert np.array_equal(out[:, 0], rec[:, 0])
    # the state is bit-identical; the energy sums differ only in reduction order
    assert np.max(np.abs(out[:, 1:] - rec[:, 1:])) <= 1e-12 * rec[:, 1].max()
    o = OracleEngine(sc.params)
    sc.load(o)
    o.step(3500, 0.0)
    _compare(g, o, sc, grid=True, exact=True)


@pytest.mark.parametrize("dim", [2, 3])
def test_binning_is_bit_exact(hip_lib, dim):
    """Stencil origin (the reference's integer support rule, floor(u + 1.0)
    with no contraction) equals the CPU restatement bit for bit, including
    particles placed one ulp below node and half-cell positions."""
    from oracle.pyoracle import OracleEngine
    bases = BASES_2D + [_ffi.BASIS_CPDI_Q4] if dim == 2 else BASES_3D
    for basis in bases:
        _say(f"binning dim={dim} basis={basis}")
        if basis == _ffi.BASIS_CPDI_Q4:
            sc = scenes.cpdi_tubes(rings=(2, 2), cells=48, nr=3, nt=48, platens=False)
        else:
            sc = (scenes.block2d(cells=16, basis=basis) if dim == 2 else scenes.block3d(nodes=(11, 11, 11), basis=basis))
            s = sc.solids[0]
            h = 1.0 / 16 if dim == 2 else 1.0 / 10
            rng = np.random.default_rng(7)
            n = s["pos"].shape[0]
            k = rng.integers(4, 8, size=(n, dim)).astype(np.float64)
            adv = k * h
            sel = rng.integers(0, 4, size=n)
            # exact nodes, one ulp below, half cells, one ulp below half cells
            adv[sel == 1] = np.nextafter(adv[sel == 1], 0.0)
            adv[sel == 2] = (k[sel == 2] + 0.5) * h
            adv[sel == 3] = np.nextafter((k[sel == 3] + 0.5) * h, 0.0)
            s["pos"] = np.ascontiguousarray(adv)
        g = Engine(sc.params)
        sc.load(g)
        o = OracleEngine(sc.params)
        sc.load(o)
        keys = g.bins()
        nx, ny = sc.params.nx, sc.params.ny
        off = 0
        for si, s in enumerate(sc.solids):
            for p in range(s["pos"].shape[0]):
                sup = o.support(si, p, True)
                if sup is None:
                    assert keys[off + p] == -1
                    continue
                near = sup[0]
                i, j = near % nx, (near // nx) % ny
                kk = near // (nx * ny)
                exp = i.min() + nx * (j.min() + ny * kk.min())
                assert keys[off + p] == exp, (basis, si, p)
            off += s["pos"].shape[0]


def test_run_to_run_bitwise_reproducible(hip_lib):
    """No floating-point atomics: two runs give identical bits."""
    sc = scenes.impact3d(nodes=(36, 36, 22), plate_mm=(16.0, 16.0, 5.0), sphere_r_mm=2.0, steps=10)
    res = []
    for _ in range(2):
        g = Engine(sc.params)
        sc.load(g)
        g.step(sc.steps, 0.0)
        res.append((g.get_solid(1), g.get_grid(), g.energies()))
        g.close()
    for k in res[0][0]:
        assert np.array_equal(res[0][0][k], res[1][0][k]), k
    for k in res[0][1]:
        assert np.array_equal(res[0][1][k], res[1][1][k]), k
    assert res[0][2] == res[1][2]


def test_error_flags_match_restatement(hip_lib):
    """A particle outside the grid, a NaN velocity: the step finishes, skips or
    propagates as the restatement does, and reports the same flags."""
    sc = scenes.block2d(cells=16, steps=3)
    s = sc.solids[0]
    s["pos"][0] = (1.5, 0.5)      # outside the grid
    s["vel"][1] = (np.nan, 0.0)   # non-finite
    from oracle.pyoracle import OracleEngine
    g = Engine(sc.params)
    sc.load(g)
    o = OracleEngine(sc.params)
    sc.load(o)
    _, rg = g.step(3, 0.0)
    _, ro = o.step(3, 0.0)
    assert rg == ro == _ffi.JUMP_ERR_NUMERIC
    fg, fo = g.flags(), o.flags()
    assert fg == fo
    assert fg & _ffi.JUMP_FLAG_OUT_OF_GRID
    assert np.array_equal(g.get_solid(0)["pos"][0], [1.5, 0.5])


def test_benchmark_size_properties(hip_lib):
    """The 3D impact at full size (4.23M particles, 256x256x64 nodes): mass is
    conserved by P2G, the grid update conserves momentum where no node is
    fixed, and the run is bitwise reproducible."""
    _say("impact3d full size")
    sc = scenes.impact3d(steps=3)
    mtot = sum(float(s["mass"].sum()) for s in sc.solids)
    ptot = sum((s["mass"][:, None] * s["vel"]).sum(0) for s in sc.solids)
    outs = []
    for rep in range(2):
        g = Engine(sc.params)
        sc.load(g)
        g.step(1, 0.0)
        gr = g.get_grid()
        assert math.isclose(gr["mass"].sum(), mtot, rel_tol=1e-12)
        # the plate and sphere lie away from the fixed faces after one step
        assert np.allclose(gr["momentum0"].sum(0), ptot, rtol=1e-9, atol=1e-9 * np.abs(ptot).max())
        g.step(2, 0.0)
        outs.append((g.get_solid(1)["pos"], g.energies()))
        g.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]


def test_impact_full_size_matches_oracle(hip_lib):
    """The 3D impact at full size (4,229,423 particles, 256x256x64 nodes, J2
    plate, quadratic B-splines, MUSL): 3 steps on the GPU equal the CPU
    restatement bit for bit, particles, material history and grid."""
    _say("impact3d full size vs oracle, 3 steps")
    sc = scenes.impact3d(steps=3)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=True)
    g.close()
    o.close()


def test_batched_256_instances(hip_lib):
    """The 256 seeded two-rings instances (seeds 1000..1255; 5.65M
    particles, grids 65 to 511 nodes wide, linear or GIMP, drawn E and speed)
    in one batch: after 4 steps every member equals its instance stepped alone
    bit for bit, energies included, and three sampled instances (the smallest,
    the largest, one GIMP) equal the CPU restatement (Neo-Hookean: 1e-10)."""
    from jump_amd.engine import Batch
    from oracle.pyoracle import OracleEngine
    _say("batched 256 instances")
    steps = 4
    scs = [scenes.two_rings_instance(1000 + i, steps=steps) for i in range(256)]
    members = []
    for sc in scs:
        m = Engine(sc.params)
        sc.load(m)
        members.append(m)
    b = Batch(members)
    b.step(steps)
    ke, se = b.energies()
    b.sync()
    _say("batched 256: stepped; comparing with each instance alone")
    for i, (m, sc) in enumerate(zip(members, scs)):
        a = Engine(sc.params)
        sc.load(a)
        a.step(steps, 0.0)
        for s in range(len(sc.solids)):
            sa, sm = a.get_solid(s), m.get_solid(s)
            for k in sa:
                assert np.array_equal(sa[k], sm[k]), (i, s, k)
        assert (ke[i], se[i]) == a.energies(), i
        a.close()
    sizes = [sc.n_particles for sc in scs]
    gimp = next(i for i, sc in enumerate(scs) if sc.params.basis == _ffi.BASIS_GIMP)
    for i in sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), gimp}):
        _say(f"batched 256: instance {i} ({sizes[i]} particles) vs oracle")
        o = OracleEngine(scs[i].params)
        scs[i].load(o)
        o.step(steps, 0.0)
        _compare(members[i], o, scs[i], grid=False, exact=False)
        o.close()
    b.close()
    for m in members:
        m.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_binned_order_matches_within_tolerance(hip_lib, dim):
    """JUMP_ORDER_BINNED sums each node bin by bin (fixed order, not the
    reference's particle order): equal to the restatement within 1e-10 over
    a short horizon, and bitwise reproducible run to run."""
    from oracle.pyoracle import OracleEngine
    if dim == 2:
        sc = scenes.block2d(cells=16, basis=_ffi.BASIS_QUAD_BSPLINE, material="j2", steps=12)
    else:
        sc = scenes.impact3d(nodes=(36, 36, 22), plate_mm=(16.0, 16.0, 5.0), sphere_r_mm=2.0, steps=25)
    sc.params.order = _ffi.ORDER_BINNED
    runs = []
    for _ in range(2):
        g = Engine(sc.params)
        sc.load(g)
        g.step(sc.steps, 0.0)
        runs.append(g)
    o = OracleEngine(sc.params)
    sc.load(o)
    o.step(sc.steps, 0.0)
    _compare(runs[0], o, sc, exact=False)
    for s in range(len(sc.solids)):
        a, b = runs[0].get_solid(s), runs[1].get_solid(s)
        for k in a:
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("dim", [2, 3])
def test_checkpoint_restart_is_bitwise(hip_lib, dim):
    """Run, checkpoint, run on; then restore into a fresh context and run the
    same steps: the two continuations agree bit for bit.  A checkpoint of
    other solids is refused."""
    if dim == 2:
        sc = scenes.block2d(cells=16, basis=_ffi.BASIS_QUAD_BSPLINE, material="j2", steps=6)
    else:
        sc = scenes.impact3d(nodes=(36, 36, 22), plate_mm=(16.0, 16.0, 5.0), sphere_r_mm=2.0, steps=12)
    a = Engine(sc.params)
    sc.load(a)
    t, _ = a.step(sc.steps, 0.0)
    blob = a.checkpoint()
    t_end, _ = a.step(sc.steps, t)
    b = Engine(sc.params)
    sc.load(b)
    b.restore(blob)
    t_b, _ = b.step(sc.steps, t)
    assert t_b == t_end
    for s in range(len(sc.solids)):
        sa, sb = a.get_solid(s), b.get_solid(s)
        for k in sa:
            assert np.array_equal(sa[k], sb[k]), k
        ga, gb = a.get_state(s), b.get_state(s)
        for k in ga:
            assert np.array_equal(ga[k], gb[k]), k
    assert a.energies() == b.energies()
    other = scenes.block2d(cells=8, steps=1) if dim == 3 else scenes.block2d(cells=12, steps=1)
    c = Engine(other.params)
    other.load(c)
    with pytest.raises(JumpError):
        c.restore(blob)


def test_batch_members_step_as_alone(hip_lib):
    """A batch of 2D instances of different grids, bases (linear, GIMP,
    quadratic B-spline), materials and dt, one with a rigid platen and a
    Dirichlet face: after stepping together, every member equals the same
    problem stepped alone, bit for bit, energies included."""
    from jump_amd.engine import Batch
    scs = [scenes.two_rings_instance(1000 + i, steps=15) for i in (1, 2, 4)]
    scs.append(scenes.rings2d(cells=24, basis=_ffi.BASIS_LINEAR, material="j2", steps=15))
    scs.append(scenes.block2d(cells=12, basis=_ffi.BASIS_QUAD_BSPLINE, material="elastic", steps=15))
    alone, members = [], []
    for sc in scs:
        a = Engine(sc.params)
        sc.load(a)
        a.step(15, 0.0)
        alone.append(a)
        m = Engine(sc.params)
        sc.load(m)
        members.append(m)
    b = Batch(members)
    b.step(10)
    b.step(5)
    ke, se = b.energies()
    b.sync()
    for i, (a, m, sc) in enumerate(zip(alone, members, scs)):
        assert b.t[i] == sum([sc.params.dt] * 15) or abs(b.t[i] - 15 * sc.params.dt) < 1e-12 * b.t[i]
        for s in range(len(sc.solids)):
            sa, sm = a.get_solid(s), m.get_solid(s)
            for k in sa:
                assert np.array_equal(sa[k], sm[k]), (i, s, k)
        assert (ke[i], se[i]) == a.energies(), i
    b.close()


def test_batch_graph_replay_and_gather_shapes_equal_eager(hip_lib):
    """HIP-graph replay (one step and chunks of 8) and every launch shape of
    the batched gather (the persistent work-queue kernel included) leave each
    member bitwise equal to the eager batch in
    the default shape, through calls of 3, 8, 1 and 12 steps with energies
    and state reads between them; the autotuner keeps the shape it measured
    fastest."""
    from jump_amd.engine import Batch
    _say("batch graphs and gather shapes")
    scs = [scenes.two_rings_instance(1000 + i, steps=24) for i in (0, 3, 5, 7)]
    scs.append(scenes.rings2d(cells=24, basis=_ffi.BASIS_LINEAR, material="j2", steps=24))
    scs.append(scenes.block2d(cells=12, basis=_ffi.BASIS_QUAD_BSPLINE, material="elastic", steps=24))

    def make():
        ms = []
        for sc in scs:
            m = Engine(sc.params)
            sc.load(m)
            ms.append(m)
        return ms, Batch(ms)

    ref_members, ref = make()
    ref.options(graph=False, autotune=False, shape=0)
    variants = [("graph+autotune", dict())] + \
        [(f"graph shape {k}", dict(graph=True, autotune=False, shape=k)) for k in range(_ffi.P2G_SHAPES)] + \
        [("eager shape 5", dict(graph=False, autotune=False, shape=5)),
         ("eager shape 8", dict(graph=False, autotune=False, shape=8))]
    runs = []
    for name, kw in variants:
        members, b = make()
        if kw:
            b.options(**kw)
        runs.append((name, members, b))
    for n in (3, 8, 1, 12):
        ref.step(n)
        kr = ref.energies()
        ref.sync()
        for name, members, b in runs:
            b.step(n)
            assert all(np.array_equal(x, y) for x, y in zip(b.energies(), kr)), (name, n)
            b.sync()
            for i, (a, m) in enumerate(zip(ref_members, members)):
                for s_ in range(len(scs[i].solids)):
                    sa, sm = a.get_solid(s_), m.get_solid(s_)
                    for k in sa:
                        assert np.array_equal(sa[k], sm[k]), (name, n, i, k)
    tun = runs[0][2].tuning()
    groups = []  # members with the same basis (and order), in order of first appearance
    for sc in scs:
        if sc.params.basis not in groups:
            groups.append(sc.params.basis)
    sizes = [sum(sc.params.basis == g for sc in scs) for g in groups]
    assert len(tun) == len(groups)
    for g, (shape, ms) in enumerate(tun):
        if sizes[g] == 1:  # a group of one member steps as a single context: nothing to tune
            assert shape == -1 and np.all(ms == 0.0)
            continue
        print("autotuner", _ffi.P2G_SHAPE_NAMES[shape], np.round(ms, 4))
        assert np.all(ms > 0.0) and ms[shape] == ms.min()
    for _, members, b in runs + [("ref", ref_members, ref)]:
        b.close()
        for m in members:
            m.close()


# ---------------------------------------------------------------- contact
CONTACT_CASES = [(b, m) for b in (_ffi.BASIS_LINEAR, _ffi.BASIS_GIMP) for m in ("elastic", "neohookean", "j2")] + \
    [(_ffi.BASIS_QUAD_BSPLINE, "elastic"), (_ffi.BASIS_CUBIC_BSPLINE, "elastic")]


@pytest.mark.parametrize("basis,material", CONTACT_CASES)
def test_contact_collision_matches_oracle(hip_lib, basis, material):
    """Two disks colliding with multi-body contact, 800 steps through the
    impact and rebound: the per-solid P2G slots, the correction and the
    per-slot G2P equal the restatement bit for bit (grid totals included;
    Neo-Hookean within 1e-10)."""
    _say(f"contact head-on basis={basis} material={material}")
    sc = scenes.disks_contact(basis=basis, material=material, steps=800)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=_exact(material))


@pytest.mark.parametrize("friction", [0.0, 0.3, 1e6])
def test_contact_oblique_friction_matches_oracle(hip_lib, friction):
    _say(f"contact oblique friction={friction}")
    sc = scenes.disks_contact(offset=0.15, friction=friction, steps=1000)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=True)


def test_contact_binned_order_and_overflow(hip_lib):
    """Binned summation order with contact (slots may fill in another order;
    same physics) within 1e-9 of the restatement; a third solid at a node
    raises the same overflow flag on both sides."""
    from oracle.pyoracle import OracleEngine
    sc = scenes.disks_contact(offset=0.15, friction=0.3, steps=600)
    sc.params.order = _ffi.ORDER_BINNED
    g = Engine(sc.params)
    sc.load(g)
    g.step(sc.steps, 0.0)
    o = OracleEngine(sc.params)
    sc.load(o)
    o.step(sc.steps, 0.0)
    for s in range(2):
        a, b = g.get_solid(s), o.get_solid(s)
        for k in ("pos", "vel", "F", "stress"):
            _close(a[k], b[k], rtol=1e-9, what=f"binned contact solid {s} {k}")
    sc = scenes.disks_contact(gap_cells=1, steps=300)
    extra = dict(sc.solids[0])
    pos = sc.solids[0]["pos"].copy()
    d = sc.solids[1]["pos"][:, 0].mean() - pos[:, 0].mean()
    pos += (0.5 * d, 0.5 * d)
    extra["pos"] = pos
    extra["vel"] = np.zeros_like(pos)
    extra["vel"][:, 1] = -0.1
    sc.solids.append(extra)
    g = Engine(sc.params)
    sc.load(g)
    o = OracleEngine(sc.params)
    sc.load(o)
    _, rg = g.step(sc.steps, 0.0)
    _, ro = o.step(sc.steps, 0.0)
    assert rg == ro == _ffi.JUMP_ERR_NUMERIC
    fg, fo = g.flags(), o.flags()  # flags are read and cleared
    assert fg == fo and fg & _ffi.JUMP_FLAG_CONTACT_OVERFLOW, (fg, fo)
    _compare(g, o, sc, exact=True)


RIGID_WALL_CASES = [(_ffi.BASIS_LINEAR, dict(fixed=True)),
                    (_ffi.BASIS_LINEAR, dict(v=(0.08, -0.1), fixed=True, friction=0.3)),
                    (_ffi.BASIS_GIMP, dict(wall_v=(0.1, 0.0), friction=1e6)),
                    (_ffi.BASIS_QUAD_BSPLINE, dict(v=(0.0, 0.0), wall_v=(0.0, 0.2), gap_cells=1)),
                    (_ffi.BASIS_CUBIC_BSPLINE, dict(fixed=True, material="j2"))]


@pytest.mark.parametrize("basis,kw", RIGID_WALL_CASES)
def test_contact_rigid_wall_matches_oracle(hip_lib, basis, kw):
    """A disk against a rigid wall with multi-body contact (the rigid body's
    own normal and velocity, the rigid body never corrected): 1000 steps
    through impact and rebound, bitwise equal to the restatement."""
    _say(f"rigid wall basis={basis} {kw}")
    sc = scenes.disk_wall(basis=basis, steps=1000, **kw)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=True)
    assert g.flags() == 0


def test_cpdi_tube_array_contact_matches_oracle(hip_lib):
    """A 3 x 3 array of CPDI rings, each its own solid, between rigid
    platens, with contact between rings and with the platens (the set-up of
    Main_LateralCompression_Cellular.jl at a small size): 200 steps equal to
    the restatement within 1e-10 (Neo-Hookean)."""
    _say("cpdi tube array 3x3 with contact")
    sc = scenes.cpdi_tubes(rings=(3, 3), cells=96, nr=2, nt=64, steps=200, contact=True)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=False)
    assert g.flags() == 0


def test_cpdi_tube_array_full_size_properties(hip_lib):
    """The CPDI tube array at full shape: an 8 x 8 array of CPDI rings (64
    solids, 1,239,784 particles with the platens) on a 2048 x 2048 grid, rigid
    platens at 1 m/s, contact between rings and platens, corners jittered by
    0.1 cell: no flag is raised, the rings' x momentum stays zero (ring-ring
    contact conserves momentum; the frictionless platens push along y only),
    the platens close the array symmetrically, and the run is bitwise
    reproducible."""
    _say("cpdi tube array 8x8 at 2048^2")
    sc = scenes.cpdi_tubes(rings=(8, 8), cells=2047, nr=16, nt=1200, steps=60, contact=True)
    runs = []
    for rep in range(2):
        g = Engine(sc.params)
        sc.load(g)
        t, rc = g.step(sc.steps, 0.0)
        assert rc == 0 and g.flags() == 0
        vs = [g.get_solid(k) for k in range(64)]
        runs.append(vs)
        g.close()
    for a, b in zip(runs[0], runs[1]):
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    m = np.concatenate([v["mass"] for v in runs[0]])
    vel = np.concatenate([v["vel"] for v in runs[0]])
    scale = float((m * np.abs(vel).sum(1)).sum())
    assert scale > 0.0
    assert abs(float((m * vel[:, 0]).sum())) <= 1e-10 * scale
    vy = np.array([float((v["mass"] * v["vel"][:, 1]).sum() / v["mass"].sum()) for v in runs[0]]).reshape(8, 8)
    assert np.all(vy[0] > 0.0) and np.all(vy[-1] < 0.0)  # bottom row pushed up, top row down


def test_contact_refusals(hip_lib):
    """Batches and the reverse mode refuse contact; 3D rigid solids do not
    exist (Solid3D has no rigid flag)."""
    from jump_amd.engine import Batch, make_material
    sc3 = scenes.spheres_contact(steps=2)
    g3 = Engine(sc3.params)
    with pytest.raises(JumpError):
        g3.add_solid(make_material(_ffi.MAT_RIGID, density=1.0), np.array([[0.5, 0.5, 0.5]]), np.zeros((1, 3)),
                     np.ones(1), np.ones(1), np.ones(1))
    g3.close()
    sc = scenes.disks_contact(steps=2)
    g = Engine(sc.params)
    sc.load(g)
    with pytest.raises(JumpError):
        Batch([g])
    from jump_amd import adjoint
    with pytest.raises(JumpError):
        adjoint.rrule(g, 2)
    g.close()


def test_cpdi_two_disks_from_gmsh(hip_lib):
    """Main_TwoDisksCPDI.jl's set-up through the host mirror: two disks of
    CPDI-Q4 domains read from disk.msh (Solid2D(fileName, mat)), placed with
    move_cpdi, on the 1.5 x 1.5 grid of 11 x 11 nodes; 300 steps of the USL
    step equal the restatement bit for bit (the example itself selects the
    2D MUSL step, whose parity test runs the same set-up)."""
    _say("cpdi two disks from disk.msh")
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
    engines = [problem.build_engine(g, [s1, s2], basis.CPDIQ4Basis(), algorithm.USL(1e-9), data, engine=cls)
               for cls in (Engine, OracleEngine)]
    for e in engines:
        _, rc = e.step(300, 0.0)
        assert rc == 0
    for s in range(2):
        a, b = engines[0].get_solid(s), engines[1].get_solid(s)
        for k in ("pos", "vel", "volume", "F", "strain", "stress"):
            _same(a[k], b[k], True, f"solid {s} {k}")
    _close(engines[0].energies(), engines[1].energies(), rtol=1e-12, what="energies")


@pytest.mark.parametrize("basis,shuffle", [(_ffi.BASIS_QUAD_BSPLINE, False), (_ffi.BASIS_LINEAR, False),
                                           (_ffi.BASIS_GIMP, False), (_ffi.BASIS_QUAD_BSPLINE, True)])
def test_run_replay_equals_merging_every_step(hip_lib, basis, shuffle):
    """The 3D passes replay recorded runs of particles wherever no particle
    entered or left a node's window since the runs were recorded.  A block
    translating at about a tenth of a cell per step (so a large share of the
    particles changes bin) is stepped 40 times in one context, and in a twin
    whose bins are read before every step, which makes every step merge at
    every node: the two agree bit for bit, grid included.  With the
    particles numbered in random order every run is one particle long and
    interior windows need more entries than a node keeps, so both passes
    merge at those nodes every step."""
    _say(f"run replay vs merging, basis={basis}, shuffle={shuffle}")
    sc = scenes.block3d(nodes=(14, 14, 14), basis=basis, v=(15.0, -20.0, 10.0), steps=40)
    if shuffle:
        s0 = sc.solids[0]
        perm = np.random.default_rng(7).permutation(s0["pos"].shape[0])
        for k, v in list(s0.items()):
            if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[0] == perm.size:
                s0[k] = v[perm]
    a = Engine(sc.params)
    sc.load(a)
    b = Engine(sc.params)
    sc.load(b)
    k0 = b.bins()
    for _ in range(sc.steps):
        a.step(1, 0.0)
        b.bins()
        b.step(1, 0.0)
    assert a.flags() == b.flags() == 0
    moved = int(np.count_nonzero(b.bins() != k0))
    assert moved > k0.size // 4, moved
    for k in ("pos", "vel", "volume", "F", "strain", "stress"):
        _same(a.get_solid(0)[k], b.get_solid(0)[k], True, k)
    ga, gb = a.get_grid(), b.get_grid()
    for k in ("mass", "momentum0", "momentum", "momentum2", "force"):
        _same(ga[k], gb[k], True, f"grid {k}")


@pytest.mark.parametrize("case", ["quad", "linear", "gimp", "quad-shuffled", "contact", "contact-wall"])
def test_run_replay_2d_equals_merging_every_step(hip_lib, case):
    """The 2D gathers (the USL step's and the contact step's) replay recorded
    runs as the 3D passes do.  A block moving 0.9 cell in 40 steps, or
    two disks in contact, or a disk against a moving rigid wall, stepped in
    one context and in a twin that reads its bins before every step (every
    step then merges at every node): bit for bit equal, grid included."""
    _say(f"2D run replay vs merging: {case}")
    if case.startswith("contact"):
        sc = (scenes.disks_contact(basis=_ffi.BASIS_GIMP, steps=300) if case == "contact" else
              scenes.disk_wall(basis=_ffi.BASIS_LINEAR, steps=300, wall_v=(0.0, 0.2), v=(0.0, 0.0), gap_cells=1))
    else:
        basis = {"quad": _ffi.BASIS_QUAD_BSPLINE, "linear": _ffi.BASIS_LINEAR,
                 "gimp": _ffi.BASIS_GIMP}[case.split("-")[0]]
        sc = scenes.block2d(cells=24, basis=basis, material="elastic", steps=40, seed=2)
        s0 = sc.solids[0]
        s0["vel"] = s0["vel"] + np.array([4.0, -3.0])
        if case.endswith("shuffled"):
            perm = np.random.default_rng(3).permutation(s0["pos"].shape[0])
            for k, v in list(s0.items()):
                if isinstance(v, np.ndarray) and v.ndim >= 1 and v.shape[0] == perm.size:
                    s0[k] = v[perm]
    a = Engine(sc.params)
    sc.load(a)
    b = Engine(sc.params)
    sc.load(b)
    k0 = b.bins()
    for _ in range(sc.steps):
        a.step(1, 0.0)
        b.bins()
        b.step(1, 0.0)
    assert a.flags() == b.flags() == 0
    moved = int(np.count_nonzero(b.bins() != k0))
    assert moved > 0, moved
    for s in range(len(sc.solids)):
        for k in ("pos", "vel", "volume", "F", "strain", "stress"):
            _same(a.get_solid(s)[k], b.get_solid(s)[k], True, f"solid {s} {k}")
    ga, gb = a.get_grid(), b.get_grid()
    for k in ("mass", "momentum0", "momentum", "force"):
        _same(ga[k], gb[k], True, f"grid {k}")


def test_usl2d_tube_cpdi_platens(hip_lib):
    """Main_LateralCompressionTube_CPDI.jl's set-up: the reference's tube.msh
    (7898 CPDI-Q4 domains, J2 steel E=210 GPa, fy=310 MPa, k=10 MPa in mm-s
    units) between rigid platens on the 100 x 60 mm grid of 51 x 21 nodes, the
    top one moving down at 10 m/s, USL(0), dt = 0.2 dx/c; 300 steps equal the
    restatement bit for bit.  The example predates the current Grid2D and
    solve_explicit_dynamics_2D signatures; the platens get unit CPDI domains
    (a rigid solid's grid coupling uses only its particles)."""
    _say("usl2d tube.msh with platens")
    from jump_amd import material, solid
    from jump_amd.engine import make_params
    rho, E, nu, fy, k1, v = 7800e-12, 210.0e3, 0.3, 310.0, 10.0, 10e3
    dx = 100.0 / 50
    f = dx / 2
    m1 = material.ElastoPlasticMaterial(E, nu, rho, fy, k1, 7898)
    tube = solid.Solid2D(os.path.join(GOLD, "tube.msh"), m1)
    sol = [dict(material=m1.to_c(), pos=tube.pos, vel=tube.velocity, mass=tube.mass, volume=tube.volume,
                volume0=tube.volumeInitial, corners=tube.nodes, elems=tube.elems)]
    for yc, vy in ((60.0 - dx / 2, -v), (dx / 2, 0.0)):
        pts = solid.buildParticleForRectangle([50.0, yc], 100.0, dx, f)
        nc = pts.shape[0]
        pc = np.concatenate([pts + [-f / 2, -f / 2], pts + [f / 2, -f / 2], pts + [f / 2, f / 2],
                             pts + [-f / 2, f / 2]])
        pe = np.stack([np.arange(nc) + k * nc for k in range(4)], axis=1)
        sol.append(dict(material=material.RigidMaterial(0.0, vy, rho).to_c(), pos=pts, vel=np.zeros_like(pts),
                        mass=np.full(nc, rho * f * f), volume=np.full(nc, f * f), volume0=np.full(nc, f * f),
                        corners=pc, elems=pe))
    dt = 0.2 * dx / math.sqrt((m1.lambda_ + 2 * m1.mu) / m1.density)
    p = make_params(2, _ffi.BASIS_CPDI_Q4, _ffi.ALG_USL, dt, (0.0, 0.0), (100.0, 60.0), (51, 21), tolerance=0.0)
    sc = scenes.Scene("tube_cpdi", p, sol, steps=300)
    g, o = _run(sc, sc.steps, hip_lib)
    assert o.get_state(0)["pstrain"].max() > 1e-3  # the platen has yielded the tube
    _compare(g, o, sc, exact=True)


CONTACT3D_CASES = [(_ffi.BASIS_QUAD_BSPLINE, m) for m in ("elastic", "neohookean", "j2")] + \
    [(_ffi.BASIS_LINEAR, "elastic"), (_ffi.BASIS_GIMP, "elastic"), (_ffi.BASIS_CUBIC_BSPLINE, "elastic")]


@pytest.mark.parametrize("basis,material", CONTACT3D_CASES)
def test_contact3d_collision_matches_oracle(hip_lib, basis, material):
    """Two spheres colliding with multi-body contact on the 3D MUSL step, 300
    steps into the impact: slot P2G, the normal correction, G2P-1, the slot
    re-projection and G2P-2 equal the restatement bit for bit (grid sums
    included; Neo-Hookean within 1e-10)."""
    _say(f"contact3d basis={basis} material={material}")
    sc = scenes.spheres_contact(gap_cells=2, basis=basis, material=material, steps=300)
    g, o = _run(sc, sc.steps, hip_lib)
    _compare(g, o, sc, exact=_exact(material))


def test_contact3d_oblique_binned_order(hip_lib):
    """An oblique 3D impact in binned summation order: within 1e-9 of the
    restatement after 400 steps."""
    from oracle.pyoracle import OracleEngine
    sc = scenes.spheres_contact(gap_cells=2, offset=0.15, steps=400)
    sc.params.order = _ffi.ORDER_BINNED
    g = Engine(sc.params)
    sc.load(g)
    g.step(sc.steps, 0.0)
    o = OracleEngine(sc.params)
    sc.load(o)
    o.step(sc.steps, 0.0)
    for s in range(2):
        a, b = g.get_solid(s), o.get_solid(s)
        for k in ("pos", "vel", "F", "stress"):
            _close(a[k], b[k], rtol=1e-9, what=f"binned contact3d solid {s} {k}")
