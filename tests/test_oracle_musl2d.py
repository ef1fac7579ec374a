"""Pins the oracle's 2D MUSL step with the blended mass matrix
(SolveExplicitSolidMechanics2D.jl:25-329) by what the mathematics fixes,
not by retyping its formulas:

- B = 0.1 diag(m) + 0.9 M maps the constant field 1 to m (the shape functions
  sum to one), so a uniformly translating stress-free body keeps its
  velocity and moves by dt v per step, and a uniform body force adds dt g to
  every particle velocity, for every basis and any FLIP fraction;
- B x = p0 is the normal equation of the least-squares problem
  min 0.9 sum_p m_p (N_p . x - v_p)^2 + 0.1 sum_I m_I (x_I - p0_I / m_I)^2,
  which the test solves with numpy's SVD-based lstsq from textbook hat
  functions: with alpha = 0 the particle velocities after one step are the
  solution read back at the particles (a swapped blend, a lumped or a purely
  consistent mass would each miss it);
- the Dirichlet mask zeroes only the nodal momenta, which neither G2P pass
  reads (:184-199), so fixing every node leaves the particles as they are;
- unsupported combinations are refused."""
import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.engine import JumpError, make_material, make_params
from oracle.pyoracle import OracleEngine

BASES = [_ffi.BASIS_LINEAR, _ffi.BASIS_QUAD_BSPLINE, _ffi.BASIS_CUBIC_BSPLINE, _ffi.BASIS_GIMP]


def _block(basis, vel, alpha, body=(0.0, 0.0), cells=12, jitter=0.0, seed=1):
    """A stress-free elastic block in the middle of the unit square."""
    rng = np.random.default_rng(seed)
    dx = 1.0 / cells
    f = dx / 2
    xs = np.arange(0.3 + f / 2, 0.7, f)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    pos = np.stack([X.ravel(), Y.ravel()], 1) + jitter * dx * rng.uniform(-1, 1, size=(X.size, 2))
    n = pos.shape[0]
    vol = np.full(n, f * f)
    mat = make_material(_ffi.MAT_ELASTIC, E=1.0e6, nu=0.3, density=1000.0)
    p = make_params(2, basis, _ffi.ALG_MUSL, 1e-4, (0.0, 0.0), (1.0, 1.0), (cells + 1, cells + 1), alpha=alpha,
                    body=body)
    o = OracleEngine(p)
    v = np.broadcast_to(np.asarray(vel, dtype=np.float64), (n, 2)).copy() if np.ndim(vel) == 1 else vel
    o.add_solid(mat, pos, v, 1000.0 * vol, vol, vol, lp=0.25 * f if basis == _ffi.BASIS_GIMP else None)
    return o, pos, v


@pytest.mark.parametrize("basis", BASES)
@pytest.mark.parametrize("alpha", [1.0, 0.5, 0.0])
def test_uniform_translation_is_exact(basis, alpha):
    v = np.array([0.37, -0.21])
    o, pos, _ = _block(basis, v, alpha, jitter=0.2)
    dt = 1e-4
    o.step(5, 0.0)
    a = o.get_solid(0)
    assert np.max(np.abs(a["vel"] - v)) <= 1e-13
    assert np.max(np.abs(a["pos"] - (pos + 5 * dt * v))) <= 1e-13
    assert np.max(np.abs(a["stress"])) <= 1e-6  # L ~ roundoff; E dt L ~ 1e-9
    assert o.flags() == 0


@pytest.mark.parametrize("basis", BASES)
def test_uniform_body_force_adds_dt_g(basis):
    g = np.array([0.0, -9.81])
    v = np.array([0.1, 0.2])
    o, pos, _ = _block(basis, v, 0.7, body=tuple(g))
    dt = 1e-4
    o.step(1, 0.0)
    a = o.get_solid(0)
    assert np.max(np.abs(a["vel"] - (v + dt * g))) <= 1e-13
    assert np.max(np.abs(a["pos"] - (pos + dt * (v + dt * g)))) <= 1e-13


def _hat(x, y, nx, ny, dx, dy):
    """Textbook bilinear hat functions of every node at (x, y): (nn,)."""
    xi = np.arange(nx) * dx
    yi = np.arange(ny) * dy
    wx = np.maximum(0.0, 1.0 - np.abs(x - xi) / dx)
    wy = np.maximum(0.0, 1.0 - np.abs(y - yi) / dy)
    return (wy[:, None] * wx[None, :]).ravel()  # node index i + nx j


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_nodal_velocity_is_the_blended_least_squares_fit(seed):
    rng = np.random.default_rng(seed)
    cells = 4
    nx = ny = cells + 1
    dx = dy = 1.0 / cells
    n = 9
    pos = rng.uniform(0.3, 0.7, size=(n, 2))
    vel = rng.normal(size=(n, 2))
    mass = rng.uniform(0.5, 2.0, size=n)
    vol = mass / 1000.0
    mat = make_material(_ffi.MAT_ELASTIC, E=1.0e6, nu=0.3, density=1000.0)
    p = make_params(2, _ffi.BASIS_LINEAR, _ffi.ALG_MUSL, 1e-4, (0.0, 0.0), (1.0, 1.0), (nx, ny), alpha=0.0)
    o = OracleEngine(p)
    o.add_solid(mat, pos, vel, mass, vol, vol)
    o.step(1, 0.0)
    a = o.get_solid(0)
    N = np.stack([_hat(x, y, nx, ny, dx, dy) for x, y in pos])  # (n, nn)
    mI = N.T @ mass
    p0 = N.T @ (mass[:, None] * vel)
    act = np.nonzero(mI > 0)[0]
    A = np.vstack([np.sqrt(0.9 * mass)[:, None] * N[:, act], np.diag(np.sqrt(0.1 * mI[act]))])
    xs = []
    for c in range(2):
        rhs = np.concatenate([np.sqrt(0.9 * mass) * vel[:, c], np.sqrt(0.1 * mI[act]) * (p0[act, c] / mI[act])])
        xs.append(np.linalg.lstsq(A, rhs, rcond=None)[0])
    fit = np.stack(xs, 1)  # nodal velocities of the active nodes
    want = N[:, act] @ fit
    scale = np.max(np.abs(vel))
    assert np.max(np.abs(a["vel"] - want)) <= 1e-11 * scale
    assert np.max(np.abs(a["pos"] - (pos + 1e-4 * want))) <= 1e-12
    # the grid keeps the momenta and masses of the P2G pass
    gr = o.get_grid()
    assert np.allclose(gr["mass"], mI, rtol=1e-13, atol=0)
    # and the fit differs from the lumped projection: the test can tell them apart
    lumped = N[:, act] @ (p0[act] / mI[act, None])
    assert np.max(np.abs(lumped - want)) > 1e-3 * scale


def test_dirichlet_mask_moves_no_particle():
    sc = scenes.block2d(cells=12, basis=_ffi.BASIS_QUAD_BSPLINE, steps=4, algorithm=_ffi.ALG_MUSL, alpha=0.9)
    a = OracleEngine(sc.params)
    sc.load(a)
    b = OracleEngine(sc.params)
    sc.fixed = np.full_like(sc.fixed, 1)
    sc.load(b)
    a.step(4, 0.0)
    b.step(4, 0.0)
    for k in ("pos", "vel", "F", "stress"):
        assert np.array_equal(a.get_solid(0)[k], b.get_solid(0)[k]), k
    gb = b.get_grid()
    assert not np.any(gb["momentum0"]) and not np.any(gb["momentum"])
    assert np.any(a.get_grid()["momentum0"])


def test_refusals():
    p = make_params(2, _ffi.BASIS_LINEAR, _ffi.ALG_MUSL, 1e-4, (0.0, 0.0), (1.0, 1.0), (11, 11), contact=True)
    with pytest.raises(JumpError):
        OracleEngine(p)
    p = make_params(2, _ffi.BASIS_LINEAR, _ffi.ALG_MUSL, 1e-4, (0.0, 0.0), (1.0, 1.0), (11, 11))
    o = OracleEngine(p)
    rig = make_material(_ffi.MAT_RIGID, density=1000.0, vx=0.0, vy=0.0)
    with pytest.raises(JumpError):
        o.add_solid(rig, np.array([[0.5, 0.5]]), np.zeros((1, 2)), np.ones(1), np.ones(1) * 1e-3)
    # the dense restatement stops at 4096 nodes
    p = make_params(2, _ffi.BASIS_LINEAR, _ffi.ALG_MUSL, 1e-4, (0.0, 0.0), (1.0, 1.0), (65, 65))
    with pytest.raises(JumpError):
        OracleEngine(p)
