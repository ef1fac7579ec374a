"""GPU: the output path.  LAMMPS dumps written by the library's writer
thread hold exactly the state at the call (the copy is ordered before later
steps on the context's stream), in the reference's text format (checked
against tests/lammps_ref.py, restated from Output.jl); the host mirror of
module Problem drives the library with OvitoOutput dumps and the
EnergiesFix record of Main_TwoDisks.jl."""
import os

import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.engine import Engine, JumpError
from lammps_ref import dump_text

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _full(st, dim):
    """Compressed symmetric stress rows (n, d*d column-major) -> (n, d, d)."""
    return np.asarray(st).reshape(-1, dim, dim).transpose(0, 2, 1)


def _expected(eng, sc, outs, lims):
    dim = eng.dim
    sols = []
    for s, sol in enumerate(sc.solids):
        a, st = eng.get_solid(s), eng.get_state(s)
        kind = sol["material"].kind
        el = kind == _ffi.MAT_ELASTIC or (dim == 2 and kind == _ffi.MAT_RIGID)
        sols.append(dict(pos=a["pos"], vel=a["vel"], stress=_full(a["stress"], dim), alpha=st["alpha"],
                         vm=st["vm"], elastic=el))
    return dump_text(dim, lims, sols, outs)


@pytest.mark.parametrize("dim", [2, 3])
def test_dump_holds_the_state_at_the_call(hip_lib, tmp_path, dim):
    print(f"dump dim={dim}", flush=True)
    if dim == 2:
        sc = scenes.rings2d(cells=40, basis=_ffi.BASIS_QUAD_BSPLINE, material="j2", steps=8)
        outs = ["pstrain", "vonMises", "sigmaxx", "sigmayy", "pressure", "vx", "vy"]
        lims = (1.0, 1.0, 0.0)
    else:
        sc = scenes.block3d(nodes=(11, 11, 11), basis=_ffi.BASIS_QUAD_BSPLINE, material="j2", steps=8)
        outs = ["pstrain", "vonMises", "vx", "vy", "vz", "pressure"]
        p = sc.params
        lims = (p.xmax - p.xmin, p.ymax - p.ymin, p.zmax - p.zmin)
    a = Engine(sc.params)
    sc.load(a)
    f3, f8 = str(tmp_path / "dump_p.3.LAMMPS"), str(tmp_path / "dump_p.8.LAMMPS")
    a.step(3, 0.0)
    a.dump_lammps(f3, outs)  # queued; the steps below run while it is written
    a.step(5, 0.0)
    a.dump_lammps(f8, outs)
    a.dump_lammps(f8, outs)  # appended, as the reference opens with "a"
    a.dump_wait()
    b = Engine(sc.params)  # the same run, read back synchronously
    sc.load(b)
    b.step(3, 0.0)
    want3 = _expected(b, sc, outs, lims)
    b.step(5, 0.0)
    want8 = _expected(b, sc, outs, lims)
    assert open(f3).read() == want3
    assert open(f8).read() == want8 + want8
    with pytest.raises(JumpError):
        a.dump_lammps(str(tmp_path / "x.LAMMPS"), ["damage"])
    a.dump_lammps(str(tmp_path / "no_such_dir" / "x.LAMMPS"), [])
    with pytest.raises(JumpError):
        a.dump_wait()
    a.close()
    b.close()


def test_problem_two_disks_with_output(hip_lib, tmp_path):
    """Main_TwoDisks.jl through the host mirror of module Problem: energies
    every 10 steps equal the committed record (1e-12), and the LAMMPS dumps
    at counters 0, 10, ... hold the state after counter + 1 steps."""
    print("problem two disks", flush=True)
    from jump_amd import algorithm, basis, bodyforce, fix, grid, material, output, problem, solid
    g = grid.Grid2D(0.0, 1.0, 0.0, 1.0, 21, 21)
    dx = 1.0 / 20
    f = dx / 2
    sols = []
    for c, v in (((0.2, 0.2), 0.1), ((0.8, 0.8), -0.1)):
        pts = solid.buildParticleForCircle(c, 0.2, f)
        s = solid.Solid2D(pts, material.ElasticMaterial(1000.0, 0.3, 1000.0, 0, 0, len(pts)))
        s.mass *= f * f  # Main_TwoDisks.jl:69-75 (mass starts at the density)
        s.volume[:] = f * f
        s.volumeInitial[:] = f * f
        s.velocity[:] = (v, v)
        sols.append(s)
    efile = str(tmp_path / "energies.txt")
    fixes = fix.EnergiesFix(sols, efile)
    out = output.OvitoOutput(10, str(tmp_path) + "/", ["vx", "vy"])
    data = {"total_time": 0.05, "dt": 1e-3, "time": 0.0, "bodyforce": bodyforce.ConstantBodyForce2D((0.0, 0.0))}
    problem.solve_explicit_dynamics_2D(g, sols, basis.LinearBasis(), algorithm.USL(1e-9), out, fixes, data)
    rec = np.load(os.path.join(GOLD, "two_disks_energy.npz"))["rec"]
    lines = open(efile).read().splitlines()
    got = np.array([[float(w) for w in ln.split()] for ln in lines])
    n = len(got)
    assert n == 5
    assert np.array_equal(got[:, 0], rec[:n, 0])
    assert np.max(np.abs(got[:, 1:] - rec[:n, 1:])) <= 1e-12 * rec[:, 1].max()
    sc = scenes.two_disks()
    e = Engine(sc.params)
    sc.load(e)
    done = 0
    for c in (0, 10, 20, 30, 40):
        e.step(c + 1 - done, 0.0)
        done = c + 1
        want = _expected(e, sc, ["vx", "vy"], (1.0, 1.0, 0.0))
        assert open(os.path.join(str(tmp_path), f"dump_p.{c}.LAMMPS")).read() == want, c
    e.close()


def test_dump_failure_is_reported_after_buffer_reuse(hip_lib, tmp_path):
    """Three dumps to an unwritable path in a row: the third waits for a
    free buffer set, which consumes the writer's failure flag; the failure
    must still reach jump_dump_wait."""
    print("dump failure sticky", flush=True)
    sc = scenes.rings2d(cells=20, basis=_ffi.BASIS_LINEAR, material="elastic", steps=1)
    a = Engine(sc.params)
    sc.load(a)
    bad = str(tmp_path / "no_such_dir" / "x.LAMMPS")
    for _ in range(3):
        a.dump_lammps(bad, ["vx"])
    with pytest.raises(JumpError) as ei:
        a.dump_wait()
    assert ei.value.code == _ffi.JUMP_ERR_IO
    good = str(tmp_path / "ok.LAMMPS")
    a.dump_lammps(good, ["vx"])
    a.dump_wait()  # the failure was reported once; a good dump succeeds again
    assert os.path.getsize(good) > 0
    a.close()
