"""The dual-number (ForwardDiff-style) instantiation of the CPU restatement,
which the reverse mode of the HIP library is checked against, pinned by
central differences of the double restatement: d L / d(E, nu, fy, k1, v0)
for a random linear functional L of the final positions and velocities,
3D MUSL with GIMP and J2 plasticity and 2D USL with the linear hat and
Neo-Hookean.  The dual run's values are those of the double run, so the
step is the same; only the derivatives are new."""
import numpy as np
import pytest

from jump_amd import _ffi, scenes
from oracle.pyoracle import OracleEngine


def _loss(sc, nsteps, pb, vb, **dp):
    m = sc.solids[0]["material"]
    keep = (m.E, m.nu, m.fy, m.k1)
    m.E += dp.get("E", 0.0)
    m.nu += dp.get("nu", 0.0)
    m.fy += dp.get("fy", 0.0)
    m.k1 += dp.get("k1", 0.0)
    v0 = sc.solids[0]["vel"].copy()
    if "v" in dp:
        sc.solids[0]["vel"] = v0 + dp["v"]
    o = OracleEngine(sc.params)
    sc.load(o)
    o.step(nsteps, 0.0)
    x = np.concatenate([o.get_solid(s)["pos"] for s in range(len(sc.solids))])
    v = np.concatenate([o.get_solid(s)["vel"] for s in range(len(sc.solids))])
    m.E, m.nu, m.fy, m.k1 = keep
    sc.solids[0]["vel"] = v0
    return float((pb * x).sum() + (vb * v).sum())


CASES = [("3d_gimp_j2", lambda: scenes.block3d(nodes=(9, 9, 9), basis=_ffi.BASIS_GIMP, material="j2", steps=4), 4),
         ("3d_linear_nh", lambda: scenes.block3d(nodes=(9, 9, 9), basis=_ffi.BASIS_LINEAR, material="neohookean",
                                                 steps=4), 4),
         ("2d_linear_nh", lambda: scenes.block2d(cells=12, basis=_ffi.BASIS_LINEAR, material="neohookean", steps=6), 6)]


@pytest.mark.parametrize("name,make,nsteps", CASES)
def test_dual_gradient_matches_central_differences(name, make, nsteps):
    sc = make()
    rng = np.random.default_rng(1)
    d, n = sc.params.dim, sc.n_particles
    pb, vb = rng.normal(size=(n, d)), rng.normal(size=(n, d))
    o = OracleEngine(sc.params)
    sc.load(o)
    g = o.dual_gradient(0, nsteps, pb, vb)
    m = sc.solids[0]["material"]
    checks = [("E", 1e-6 * m.E, g["E"], 1e-6), ("nu", 1e-6, g["nu"], 1e-6)]
    if m.fy:
        checks += [("fy", 1e-6 * m.fy, g["fy"], 1e-6), ("k1", 1e-6 * m.k1, g["k1"], 1e-4)]
    else:
        assert g["fy"] == 0.0 and g["k1"] == 0.0  # elastic: no dependence
    for key, h, got, rtol in checks:
        fd = (_loss(sc, nsteps, pb, vb, **{key: h}) - _loss(sc, nsteps, pb, vb, **{key: -h})) / (2 * h)
        assert abs(got - fd) <= rtol * abs(fd), (name, key, got, fd)
    for k in range(d):
        dv = np.zeros(d)
        dv[k] = 1e-6
        fd = (_loss(sc, nsteps, pb, vb, v=dv) - _loss(sc, nsteps, pb, vb, v=-dv)) / 2e-6
        assert abs(g["v0"][k] - fd) <= 1e-6 * abs(fd), (name, k, g["v0"][k], fd)
