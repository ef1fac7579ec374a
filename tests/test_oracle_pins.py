"""Pins of the CPU restatement's kinematics, return maps, energies and rigid
override by closed forms and identities that do not retype its formulas;
each check is chosen so that a plausible slip (a transposed velocity
gradient, a swapped FLIP/PIC weight, a dropped factor 2 on an engineering
shear strain, a rate ratio taken after its update, a missing 1/2, an
override applied to one momentum only) fails it.

  velocity gradient and F      an affine velocity field v = A x + c on a
    (2D L[a,b] = v_a d_b N,      symmetric particle lattice: interior nodes
     3D L[a,b] = d_a N v_b)      carry v_I = A x_I + c exactly, so the 2D USL
                                 step must give F = I + A dt and the 3D MUSL
                                 step F = I + A^T dt (SolveExplicitSolid-
                                 Mechanics2D.jl:515-523, 3D.jl:189-197)
  MUSL blend (3D.jl:109-140)     zero stress: FLIP (alpha=1) keeps v bitwise;
                                 PIC (alpha=0) gives the nodal velocities
                                 interpolated; alpha=1/2 their mean; the
                                 re-projection is sum_p N m v_new
  2D J2 (Material.jl:174-225)    after every return: stress = elastic law of
                                 (total - plastic) strain with engineering
                                 shear, and ||dev s|| = fy + k1 alpha
  Johnson-Cook (:386-441)        the alphadot filter's geometric decay over
                                 elastic calls; the rate-dependent yield
                                 surface (A + B a^n)(1 + ratio)^C reached by
                                 the radial return; the return's plastic
                                 increment (J2_trial - ys)/(3G)
  strain energy (Fix.jl:263)     SE = V/2 (lambda tr(eps)^2 + 2 mu eps:eps)
                                 for a uniform Hooke strain, KE = M |v|^2/2
  rigid override (2D.jl:448-479) every node of a rigid particle's cell holds
                                 p = p0 = m v_rigid in the rigid velocity's
                                 nonzero components (0 if fixed) and the
                                 deformable step's values elsewhere
"""
import math

import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.engine import make_material, make_params
from oracle import pyoracle
from oracle.pyoracle import OracleEngine


def _lattice(dim, cells_lo, cells_hi, dx):
    """Particles at (i + 1/4) dx and (i + 3/4) dx of cells cells_lo..cells_hi-1
    on every axis: the layout around every interior node is symmetric."""
    offs = np.concatenate([(np.arange(cells_lo, cells_hi) + q) * dx for q in (0.25, 0.75)])
    offs.sort()
    g = np.meshgrid(*([offs] * dim), indexing="ij")
    return np.stack([a.ravel() for a in g], axis=1)


AFFINE_CASES = [(2, _ffi.BASIS_LINEAR), (2, _ffi.BASIS_QUAD_BSPLINE), (2, _ffi.BASIS_CUBIC_BSPLINE),
                (3, _ffi.BASIS_LINEAR), (3, _ffi.BASIS_QUAD_BSPLINE)]


@pytest.mark.parametrize("dim,basis", AFFINE_CASES)
def test_affine_velocity_field_gives_the_velocity_gradient(dim, basis):
    cells = 16 if dim == 2 else 20
    dx = 1.0 / cells
    lo, hi = 2, cells - 2
    pos = _lattice(dim, lo, hi, dx)
    A = np.array([[0.3, -1.1, 0.4], [0.7, 0.2, -0.5], [-0.9, 0.6, 0.1]])[:dim, :dim]
    c = np.array([0.05, -0.02, 0.03])[:dim]
    vel = pos @ A.T + c
    rho = 1000.0
    vol = (0.5 * dx) ** dim
    dt = 1e-3
    alg = _ffi.ALG_USL if dim == 2 else _ffi.ALG_MUSL
    p = make_params(dim, basis, alg, dt, (0.0,) * dim, (1.0,) * dim, (cells + 1,) * dim, tolerance=1e-12,
                    alpha=0.7)
    eng = OracleEngine(p)
    n = pos.shape[0]
    eng.add_solid(scenes.elastic(1.0e3, 0.3, rho), pos, vel, np.full(n, rho * vol), np.full(n, vol))
    eng.step(1, 0.0)
    r = eng.get_solid(0)
    # particles whose support nodes see a symmetric lattice; the 3D MUSL
    # step projects twice (momentum, then momentum2 from the new velocities),
    # which doubles the margin the quadratic B-splines need
    margin = 2 if basis == _ffi.BASIS_LINEAR else (3 if dim == 2 else 6)
    inner = np.all((pos > (lo + margin) * dx) & (pos < (hi - margin) * dx), axis=1)
    assert inner.sum() >= 16
    F = r["F"][inner].reshape(-1, dim, dim).transpose(0, 2, 1)  # column-major rows -> matrices
    Lexp = A if dim == 2 else A.T  # the 3D step builds the transposed gradient
    want = np.eye(dim) + Lexp * dt
    err = np.max(np.abs(F - want))
    assert err <= 1e-9 * np.abs(A).max() * dt, err
    other = np.eye(dim) + (A.T if dim == 2 else A) * dt
    assert np.max(np.abs(F - other)) > 0.1 * np.abs(A - A.T).max() * dt  # the test tells the two apart
    # x += dt sum N v_I = dt (A x_p + c) (linear reproduction)
    dxp = r["pos"][inner] - pos[inner]
    assert np.max(np.abs(dxp - dt * vel[inner])) <= 1e-9 * dt * np.abs(vel).max()
    # strain += dt sym(L)
    eps = r["strain"][inner].reshape(-1, dim, dim)
    assert np.max(np.abs(eps - 0.5 * (A + A.T) * dt)) <= 1e-9 * np.abs(A).max() * dt


def _block3d_random(alpha, seed=4):
    cells = 10
    dx = 1.0 / cells
    pos = _lattice(3, 2, 8, dx) + 0.013 * dx  # off the symmetric layout: a general field
    rng = np.random.default_rng(seed)
    vel = rng.normal(size=pos.shape)
    rho, vol = 1000.0, (0.5 * dx) ** 3
    p = make_params(3, _ffi.BASIS_QUAD_BSPLINE, _ffi.ALG_MUSL, 1e-3, (0.0,) * 3, (1.0,) * 3, (cells + 1,) * 3,
                    alpha=alpha)
    eng = OracleEngine(p)
    n = pos.shape[0]
    eng.add_solid(scenes.elastic(1.0e3, 0.3, rho), pos, vel, np.full(n, rho * vol), np.full(n, vol))
    return eng, pos, vel, np.full(n, rho * vol)


@pytest.mark.parametrize("alpha", [1.0, 0.0, 0.5])
def test_musl_blend_and_reprojection(alpha):
    eng, pos, vel, mass = _block3d_random(alpha)
    n = pos.shape[0]
    # the weights at the start-of-step positions (G2P-1 reads them)
    sup0 = [eng.support(0, p, grad=False) for p in range(n)]
    eng.step(1, 0.0)
    g = eng.get_grid()
    vnew = eng.get_solid(0)["vel"]
    # zero stress, no body force: momentum == momentum0 at every node
    assert np.array_equal(g["momentum"], g["momentum0"])
    pic = np.zeros((n, 3))
    for p, (near, N, _) in enumerate(sup0):
        m = g["mass"][near]
        ok = m > 0
        pic[p] = ((N[ok] / m[ok])[:, None] * g["momentum"][near[ok]]).sum(0)
    if alpha == 1.0:
        assert np.array_equal(vnew, vel)  # FLIP with no force changes nothing, bit for bit
    want = alpha * vel + (1.0 - alpha) * pic
    assert np.max(np.abs(vnew - want)) <= 1e-12 * np.abs(vel).max()
    if alpha == 0.0:
        assert np.max(np.abs(vnew - vel)) > 1e-3  # PIC smooths a random field: the cases differ
    # momentum2 = sum_p N m v_new
    p2 = np.zeros_like(g["momentum2"])
    for p, (near, N, _) in enumerate(sup0):
        np.add.at(p2, near, (N * mass[p])[:, None] * vnew[p])
    assert np.max(np.abs(p2 - g["momentum2"])) <= 1e-12 * np.abs(p2).max()


def _mat(kind, **kw):
    return make_material(kind, **kw)


def test_j2_2d_return_is_elastic_law_of_elastic_strain():
    E, nu, fy, k1 = 200.0, 0.3, 0.05, 20.0
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    kappa = lam + mu  # Material.jl:161
    m = _mat(_ffi.MAT_ELASTOPLASTIC, E=E, nu=nu, density=1.0, fy=fy, k1=k1)
    state = None
    yielded = 0
    alpha_prev = 0.0
    for k in range(1, 50):
        # shear-dominated path that turns half way, with some volume change
        g12 = 2e-4 * k if k < 25 else 2e-4 * 25 - 1.5e-4 * (k - 25)
        eps = np.array([[1.2e-4 * k, g12], [g12, -0.4e-4 * k]])
        sig, state = pyoracle.stress(2, m, eps.T.reshape(-1), state=state)
        S = sig.reshape(2, 2).T
        alpha, ep = state[0], state[3:6]  # ep = [ep11, ep22, engineering ep12]
        tr = eps[0, 0] + eps[1, 1]
        kt = kappa * tr
        assert math.isclose(0.5 * (S[0, 0] + S[1, 1]), kt, rel_tol=1e-12, abs_tol=1e-15)
        dev = np.array([S[0, 0] - kt, S[1, 1] - kt, S[0, 1]])
        # elastic law on the elastic strain (2D deviator of the reference's
        # I_dev; engineering shear gamma = 2 eps12, sigma12 = mu (gamma - gamma_p))
        edev = np.array([0.5 * (eps[0, 0] - eps[1, 1]), 0.5 * (eps[1, 1] - eps[0, 0]), 2.0 * eps[0, 1]])
        law = np.array([2 * mu * (edev[0] - ep[0]), 2 * mu * (edev[1] - ep[1]), mu * (edev[2] - ep[2])])
        assert np.allclose(dev, law, rtol=1e-11, atol=1e-14 * E), (k, dev, law)
        nrm = math.sqrt(dev[0] ** 2 + dev[1] ** 2 + 2 * dev[2] ** 2)
        if alpha > alpha_prev:
            yielded += 1
            assert math.isclose(nrm, fy + k1 * alpha, rel_tol=1e-11)
            assert abs(ep[2]) > 0.0
        else:
            assert nrm <= fy + k1 * alpha + 1e-14
        alpha_prev = alpha
    assert yielded >= 10


def _jc(C, n, epsdot0=1e-3, A=0.3, B=1.0, E=200.0, nu=0.3, rho=2.0, cell=0.5):
    return _mat(_ffi.MAT_JOHNSONCOOK, E=E, nu=nu, density=rho, A=A, B=B, C=C, n=n, epsdot0=epsdot0, cellsize=cell)


def _jc_tav(E=200.0, nu=0.3, rho=2.0, cell=0.5):
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return 1000.0 * cell / math.sqrt((lam + 2 * mu) / rho), mu  # relaxation time of the rate filter


def test_johnson_cook_alphadot_filter_decays_geometrically():
    m = _jc(C=0.2, n=0.5)
    tav, _ = _jc_tav()
    dt = 7e-3
    a0 = 0.04
    state = np.zeros(12)
    state[0], state[1] = 0.01, a0
    sig = np.zeros(9)
    for k in range(1, 6):
        sig, state = pyoracle.stress(3, m, np.zeros(9), deps=np.zeros(9), dt=dt, sig=sig, state=state)
        assert state[0] == 0.01  # elastic: no plastic increment
        assert math.isclose(state[1], a0 * (1.0 - dt / tav) ** k, rel_tol=1e-13)


@pytest.mark.parametrize("ratio", [250.0, 0.3])
def test_johnson_cook_rate_dependent_yield_surface(ratio):
    C, nexp, epsdot0, A, B = 0.15, 0.5, 1e-3, 0.3, 1.0
    m = _jc(C=C, n=nexp, epsdot0=epsdot0, A=A, B=B)
    tav, mu = _jc_tav()
    a_old = 0.02
    ad_old = ratio * epsdot0
    dt = 1e-3
    deps = np.zeros((3, 3))
    deps[0, 1] = deps[1, 0] = 4e-3
    deps[2, 2] = 1e-3
    sig0 = np.diag([0.05, -0.02, 0.01])
    state = np.zeros(12)
    state[0], state[1] = a_old, ad_old
    sig, state = pyoracle.stress(3, m, deps.T.reshape(-1), deps=deps.T.reshape(-1), dt=dt,
                                 sig=sig0.T.reshape(-1), state=state)
    S = sig.reshape(3, 3).T
    s = S - np.trace(S) / 3 * np.eye(3)
    J2 = math.sqrt(1.5) * np.linalg.norm(s)
    # the yield stress at the incoming rate ratio (clamped below at 1)
    ys = (A + B * a_old ** nexp) * (1.0 + max(ratio, 1.0)) ** C
    assert math.isclose(J2, ys, rel_tol=1e-6)
    # radial return: plastic increment (J2_trial - ys)/(3G), trial = sigma + 2G deps
    tr = sig0 + 2 * mu * deps
    st = tr - np.trace(tr) / 3 * np.eye(3)
    J2t = math.sqrt(1.5) * np.linalg.norm(st)
    dp = state[0] - a_old
    assert dp > 0.0
    assert math.isclose(dp, (J2t - ys) / (3 * mu), rel_tol=1e-6)
    # the filtered rate: decay, then the increment over the relaxation time
    assert math.isclose(state[1], ad_old * (1 - dt / tav) + dp / tav, rel_tol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_strain_energy_of_a_uniform_strain(dim):
    E, nu, rho = 3.0e3, 0.27, 500.0
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    eps = np.array([[2e-3, 7e-4, -3e-4], [7e-4, -1e-3, 5e-4], [-3e-4, 5e-4, 4e-4]])[:dim, :dim]
    sig = lam * np.trace(eps) * np.eye(dim) + 2 * mu * eps
    cells = 10
    dx = 1.0 / cells
    pos = _lattice(dim, 3, 7, dx)
    n = pos.shape[0]
    vol = (0.5 * dx) ** dim * (1.0 + 0.1 * np.arange(n) / n)  # unequal volumes
    v = np.array([0.3, -0.4, 0.12])[:dim]
    alg = _ffi.ALG_USL if dim == 2 else _ffi.ALG_MUSL
    p = make_params(dim, _ffi.BASIS_LINEAR, alg, 1e-3, (0.0,) * dim, (1.0,) * dim, (cells + 1,) * dim)
    eng = OracleEngine(p)
    eng.add_solid(scenes.elastic(E, nu, rho), pos, np.tile(v, (n, 1)), rho * vol, vol,
                  strain=np.tile(eps.T.reshape(-1), (n, 1)), stress=np.tile(sig.T.reshape(-1), (n, 1)))
    ke, se = eng.energies()
    w = 0.5 * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))  # strain energy density
    assert math.isclose(se, vol.sum() * w, rel_tol=1e-12)
    assert math.isclose(ke, 0.5 * (rho * vol).sum() * float(v @ v), rel_tol=1e-12)


@pytest.mark.parametrize("vx,vy,fixed", [(0.0, -0.5, False), (0.3, -0.5, False), (0.3, 0.2, True)])
def test_rigid_override_sets_nodal_momentum(vx, vy, fixed):
    sc = scenes.rings2d(cells=40, basis=_ffi.BASIS_QUAD_BSPLINE, material="elastic", steps=1)
    sc.solids[2]["material"] = scenes.rigid(vx, vy, 7800.0, fixed=fixed)
    # move the platen down onto the rings so that some nodes carry both
    sc.solids[2]["pos"] = sc.solids[2]["pos"] - [0.0, 0.07]
    with_rigid = OracleEngine(sc.params)
    sc.load(with_rigid)
    with_rigid.step(1, 0.0)
    ga = with_rigid.get_grid()
    alone = scenes.Scene("rings", sc.params, sc.solids[:2], fixed=sc.fixed)
    ref = OracleEngine(sc.params)
    alone.load(ref)
    ref.step(1, 0.0)
    gb = ref.get_grid()
    # nodes of the rigid particles' hat cells (Basis.jl:63-78)
    nx = sc.params.nx
    dx = (sc.params.xmax - sc.params.xmin) / (nx - 1)
    x = sc.solids[2]["pos"]
    i = np.floor((x[:, 0] - sc.params.xmin) / dx).astype(int)
    j = np.floor((x[:, 1] - sc.params.ymin) / dx).astype(int)
    nodes = np.unique(np.concatenate([i + nx * j, i + 1 + nx * j, i + nx * (j + 1), i + 1 + nx * (j + 1)]))
    other = np.setdiff1d(np.arange(ga["mass"].size), nodes)
    assert np.array_equal(ga["mass"], gb["mass"])  # the rigid solid adds no mass
    for key in ("momentum", "momentum0"):
        assert np.array_equal(ga[key][other], gb[key][other])
    m = gb["mass"][nodes]
    assert (m > 0).sum() > 20  # the platen overlaps the rings' support
    for a, v in enumerate((vx, vy)):
        for key in ("momentum", "momentum0"):
            got = ga[key][nodes, a]
            if fixed:
                assert np.all(got == 0.0)
            elif v != 0.0:
                assert np.array_equal(got, m * v)
            else:
                assert np.array_equal(got, gb[key][nodes, a])
