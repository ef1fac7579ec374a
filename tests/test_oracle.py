"""Pins the CPU restatement (the oracle) with checks that do not retype its
formulas: the reference's own B-spline test script evaluated exactly, closed
forms, partition of unity and linear reproduction, finite differences of
energies, brute force on tiny inputs, conservation invariants."""
import math
import os

import numpy as np
import pytest

from jump_amd import _ffi, scenes
from jump_amd.engine import make_material, make_params
from oracle import pyoracle
from oracle.pyoracle import OracleEngine

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- B-splines
def test_quadratic_bsplines_match_reference_test_script():
    """test_quadratic_bsplines.jl:3-73 (types 1,2,3,4,1 on nodes 0..4)."""
    g = np.load(os.path.join(GOLD, "bsplines.npz"))
    types = (1, 2, 3, 4, 1)
    checked = 0
    for p, x in enumerate(g["points"]):
        for k, t in enumerate(types):
            if np.isnan(g["quad"][p, k]):
                continue
            N, _ = pyoracle.bspline(2, t, x - k)
            assert abs(N - g["quad"][p, k]) <= 4e-16, (x, k, N, g["quad"][p, k])
            checked += 1
    assert checked >= 380


def test_quadratic_bsplines_partition_of_unity_and_derivative():
    types = (1, 2, 3, 4, 1)
    for x in np.linspace(0.0, 4.0, 257):
        s = sum(pyoracle.bspline(2, t, x - k)[0] for k, t in enumerate(types))
        ds = sum(pyoracle.bspline(2, t, x - k)[1] for k, t in enumerate(types))
        assert abs(s - 1.0) < 1e-14
        if 0.0 < x < 4.0:  # the ends are kinks with one-sided derivatives
            assert abs(ds) < 1e-13
    h = 1e-6
    for t in (1, 2, 3, 4):
        for r in np.linspace(-1.49, 1.49, 119):
            if min(abs(r - b) for b in (-1.5, -1.0, -0.5, 0.0, 0.5, 1.0, 1.5)) < 2e-6:
                continue
            fd = (pyoracle.bspline(2, t, r + h)[0] - pyoracle.bspline(2, t, r - h)[0]) / (2 * h)
            assert abs(fd - pyoracle.bspline(2, t, r)[1]) < 1e-7


def test_cubic_bspline_interior_closed_form_and_boundary_partition():
    g = np.load(os.path.join(GOLD, "bsplines.npz"))
    for r, v in zip(g["cubic_r"], g["cubic"]):
        N, _ = pyoracle.bspline(3, 3, r)
        # decimal literals 0.16666666666666666 etc. in Bsplines.jl
        assert abs(N - v) < 1e-15
    # boundary-modified family sums to one over a 7-node grid
    n = 7
    for x in np.linspace(0.0, 6.0, 301):
        s = 0.0
        ds = 0.0
        for k in range(n):
            node = k + 1
            t = 1 if node in (1, n) else (2 if node == 2 else (4 if node == n - 1 else 3))
            N, dN = pyoracle.bspline(3, t, x - k)
            s += N
            ds += dN
        # type 4 writes 0.6666666667 (Bsplines.jl:168): 3e-11 from 2/3
        assert abs(s - 1.0) < 1e-10
        if 0.0 < x < 6.0:  # one-sided derivatives at the grid ends
            assert abs(ds) < 1e-10
    h = 1e-6
    for t in (1, 2, 3, 4):
        for r in np.linspace(-1.99, 1.99, 91):
            if min(abs(r - b) for b in (-2, -1, 0, 1, 2)) < 2e-6:
                continue
            fd = (pyoracle.bspline(3, t, r + h)[0] - pyoracle.bspline(3, t, r - h)[0]) / (2 * h)
            assert abs(fd - pyoracle.bspline(3, t, r)[1]) < 1e-6


# ---------------------------------------------------------------- GIMP
def _gimp_quadrature(xi, h, lp, m=20001):
    # S(xi) = (1/2lp) * integral over [xi-lp, xi+lp] of the hat (1 - |s|/h)+
    s = np.linspace(xi - lp, xi + lp, m)
    hat = np.maximum(0.0, 1.0 - np.abs(s) / h)
    return np.trapezoid(hat, s) / (2 * lp)


@pytest.mark.parametrize("lp", [0.05, 0.25, 0.5])
def test_gimp_equals_hat_convolved_with_particle(lp):
    h = 1.0
    for xi in np.linspace(-1.6, 1.6, 97):
        N, dN = pyoracle.gimp1d(xi, h, lp)
        assert abs(N - _gimp_quadrature(xi, h, lp)) < 2e-8
        e = 1e-6
        fd = (pyoracle.gimp1d(xi + e, h, lp)[0] - pyoracle.gimp1d(xi - e, h, lp)[0]) / (2 * e)
        assert abs(fd - dN) < 1e-6


def test_gimp_limits_and_partition():
    h = 0.5
    for x in np.linspace(0.3, 0.7, 41):
        N = [pyoracle.gimp1d(x - k * h, h, 1e-9)[0] for k in range(-3, 4)]
        hat = [max(0.0, 1.0 - abs(x - k * h) / h) for k in range(-3, 4)]
        assert np.allclose(N, hat, atol=1e-8)
        for lp in (0.1, 0.25):
            s = sum(pyoracle.gimp1d(x - k * h, h, lp)[0] for k in range(-3, 4))
            assert abs(s - 1.0) < 1e-14


# ---------------------------------------------------------------- supports
def _single(dim, basis, pos, nodes, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0), lp=None, corners=None, elems=None):
    p = make_params(dim, basis, _ffi.ALG_USL if dim == 2 else _ffi.ALG_MUSL, 1e-3, lo, hi, nodes)
    eng = OracleEngine(p)
    pos = np.atleast_2d(pos)
    n = pos.shape[0]
    eng.add_solid(scenes.elastic(1.0, 0.3, 1.0), pos, np.zeros_like(pos), np.ones(n), np.ones(n), lp=lp,
                  corners=corners, elems=elems)
    return eng


def _node_xyz(eng, ids, lo, hi):
    nx, ny, nz = eng.nodes
    d = eng.dim
    h = [(hi[k] - lo[k]) / (eng.nodes[k] - 1) for k in range(d)]
    i = ids % nx
    j = (ids // nx) % ny
    out = [lo[0] + i * h[0], lo[1] + j * h[1]]
    if d == 3:
        out.append(lo[2] + (ids // (nx * ny)) * h[2])
    return np.stack(out, axis=1)


@pytest.mark.parametrize("basis", [_ffi.BASIS_LINEAR, _ffi.BASIS_QUAD_BSPLINE, _ffi.BASIS_CUBIC_BSPLINE,
                                   _ffi.BASIS_GIMP])
@pytest.mark.parametrize("dim", [2, 3])
def test_partition_of_unity_and_linear_reproduction(basis, dim):
    rng = np.random.default_rng(3)
    nodes = (17, 17, 17)[:dim]
    lo, hi = (0.0,) * 3, (1.0,) * 3
    h = 1.0 / 16
    lp = 0.5 * 0.5 * h if basis == _ffi.BASIS_GIMP else None
    for _ in range(25):
        x = rng.uniform(0.25, 0.75, dim)
        eng = _single(dim, basis, x, nodes, lo, hi, lp=lp)
        for grad in (True, False):
            near, N, dN = eng.support(0, 0, grad)
            assert abs(N.sum() - 1.0) < 1e-13
            xI = _node_xyz(eng, near, lo, hi)
            assert np.allclose((N[:, None] * xI).sum(0), x, atol=1e-13)
            if grad:
                assert np.allclose(dN.sum(0), 0.0, atol=1e-10)
                # sum_I dN_I x_I = identity
                assert np.allclose(dN.T @ xI, np.eye(dim), atol=1e-10)


def test_linear_basis_stencil_is_the_cell_and_brute_force_weights():
    eng = _single(2, _ffi.BASIS_LINEAR, [0.33, 0.61], (11, 11))
    near, N, dN = eng.support(0, 0, True)
    # cell (3, 6) of a 0.1 grid, 30% / 10% into it; node order
    # [iIndex, iIndex+1, iIndex+Nx, iIndex+1+Nx] (Basis.jl:77)
    assert list(near) == [3 + 6 * 11, 4 + 6 * 11, 3 + 7 * 11, 4 + 7 * 11]
    assert np.allclose(N, [0.7 * 0.9, 0.3 * 0.9, 0.7 * 0.1, 0.3 * 0.1], atol=1e-14)
    assert np.allclose(dN[:, 0], [-9.0, 9.0, -1.0, 1.0], atol=1e-12)
    assert np.allclose(dN[:, 1], [-7.0, -3.0, 7.0, 3.0], atol=1e-12)


@pytest.mark.parametrize("n", [10, 11])
def test_qbs_getShapeFunctions_boundary_quirk(n):
    """`id == 1 | id == N` (Basis.jl:1079) is true only for id == N with N
    odd: node 1 gets type 3 there, node N type 1 only when N is odd."""
    h = 1.0 / (n - 1)
    x = np.array([0.3 * h, 0.5])  # first cell along x
    eng = _single(2, _ffi.BASIS_QUAD_BSPLINE, x, (n, n))
    near_g, Ng, _ = eng.support(0, 0, True)
    near_f, Nf, _ = eng.support(0, 0, False)
    assert list(near_g) == list(near_f)
    r = 0.3
    t1 = pyoracle.bspline(2, 1, r)[0]
    t3 = pyoracle.bspline(2, 3, r)[0]
    ix = near_g % n
    wy = Ng[ix == 0] / t1
    assert np.allclose(Nf[ix == 0], wy * t3, rtol=1e-14)
    # last cell along x
    x2 = np.array([1.0 - 0.3 * h, 0.5])
    eng2 = _single(2, _ffi.BASIS_QUAD_BSPLINE, x2, (n, n))
    near_g2, Ng2, _ = eng2.support(0, 0, True)
    _, Nf2, _ = eng2.support(0, 0, False)
    last = (near_g2 % n) == n - 1
    if n % 2 == 1:
        assert np.array_equal(Ng2[last], Nf2[last])
    else:
        assert not np.allclose(Ng2[last], Nf2[last])


def test_cpdi_rectangle_reduces_to_averaged_hat():
    """Axis-aligned rectangular domain: wf = 1/4 per corner, gradient
    weights +-1/(2a), +-1/(2b) (the textbook CPDI-Q4 reduction)."""
    a, b = 0.06, 0.04
    x0, y0 = 0.413, 0.527
    corners = np.array([[x0, y0], [x0 + a, y0], [x0 + a, y0 + b], [x0, y0 + b]])
    elems = np.array([[0, 1, 2, 3]])
    eng = _single(2, _ffi.BASIS_CPDI_Q4, corners.mean(0), (11, 11), corners=corners, elems=elems)
    near, N, dN = eng.support(0, 0, True)
    xI = _node_xyz(eng, near, (0, 0, 0), (1, 1, 1))
    hat = lambda p, q: np.maximum(0, 1 - np.abs(p - q) / 0.1)
    exp = np.array([sum(0.25 * hat(c[0], X[0]) * hat(c[1], X[1]) for c in corners) for X in xI])
    assert np.allclose(N, exp, atol=1e-14)
    assert abs(N.sum() - 1.0) < 1e-14
    # gradient: (1/V) * integral over the domain of grad of the bilinear interpolant
    S = lambda X, c: hat(c[0], X[0]) * hat(c[1], X[1])
    gx = np.array([(S(X, corners[1]) + S(X, corners[2]) - S(X, corners[0]) - S(X, corners[3])) / (2 * a) for X in xI])
    gy = np.array([(S(X, corners[2]) + S(X, corners[3]) - S(X, corners[0]) - S(X, corners[1])) / (2 * b) for X in xI])
    assert np.allclose(dN[:, 0], gx, atol=1e-10)
    assert np.allclose(dN[:, 1], gy, atol=1e-10)


# ---------------------------------------------------------------- materials
def _mat(kind, **kw):
    return make_material(kind, **kw)


def test_det_matches_numpy():
    rng = np.random.default_rng(0)
    for d in (2, 3):
        for _ in range(20):
            A = rng.normal(size=(d, d))
            assert abs(pyoracle.det(d, A.T.reshape(-1)) - np.linalg.det(A)) < 1e-13


def test_hooke_uniaxial_strain():
    E, nu = 200.0, 0.3
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    for d in (2, 3):
        eps = np.zeros((d, d))
        eps[0, 0] = 1e-3
        sig, _ = pyoracle.stress(d, _mat(_ffi.MAT_ELASTIC, E=E, nu=nu, density=1.0), eps.T.reshape(-1))
        S = sig.reshape(d, d).T
        assert math.isclose(S[0, 0], (lam + 2 * mu) * 1e-3, rel_tol=1e-14)
        assert math.isclose(S[1, 1], lam * 1e-3, rel_tol=1e-14)
        assert abs(S[0, 1]) == 0.0


def _nh_energy(F, lam, mu):
    J = np.linalg.det(F)
    d = F.shape[0]
    return 0.5 * mu * (np.trace(F.T @ F) - d) - mu * math.log(J) + 0.5 * lam * math.log(J) ** 2


@pytest.mark.parametrize("d", [2, 3])
def test_neo_hookean_is_derivative_of_its_energy(d):
    """Cauchy stress = (1/J) dW/dF F^T for W = mu/2 (I1-d) - mu lnJ + lam/2 ln^2 J,
    derivative by central differences."""
    E, nu = 10.0, 0.3
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    rng = np.random.default_rng(1)
    for _ in range(5):
        F = np.eye(d) + 0.1 * rng.normal(size=(d, d))
        J = np.linalg.det(F)
        P = np.zeros((d, d))
        h = 1e-6
        for a in range(d):
            for b in range(d):
                Fp, Fm = F.copy(), F.copy()
                Fp[a, b] += h
                Fm[a, b] -= h
                P[a, b] = (_nh_energy(Fp, lam, mu) - _nh_energy(Fm, lam, mu)) / (2 * h)
        sig_fd = P @ F.T / J
        sig, _ = pyoracle.stress(d, _mat(_ffi.MAT_NEOHOOKEAN, E=E, nu=nu, density=1.0), np.zeros(d * d),
                                 F=F.T.reshape(-1), J=J)
        assert np.allclose(sig.reshape(d, d).T, sig_fd, atol=1e-7)


def test_j2_3d_return_map_consistency():
    E, nu, fy, k1 = 200.0, 0.3, 0.2, 10.0
    mu = E / (2 * (1 + nu))
    m = _mat(_ffi.MAT_ELASTOPLASTIC, E=E, nu=nu, density=1.0, fy=fy, k1=k1)
    state = None
    eps = np.zeros((3, 3))
    alpha_prev = 0.0
    yielded = False
    for k in range(1, 60):
        eps[0, 1] = eps[1, 0] = 1e-4 * k  # pure shear path
        eps[2, 2] = -2e-5 * k
        sig, state = pyoracle.stress(3, m, eps.T.reshape(-1), state=state)
        S = sig.reshape(3, 3).T
        s = S - np.trace(S) / 3 * np.eye(3)
        ns = np.linalg.norm(s)
        alpha = state[0]
        if alpha > alpha_prev:
            yielded = True
            # consistency: ||s|| = fy + k1*alpha after the return
            assert abs(ns - (fy + k1 * alpha)) < 1e-9 * (fy + k1 * alpha) + 1e-12
        else:
            assert ns <= fy + k1 * alpha + 1e-12
        alpha_prev = alpha
    assert yielded
    # first yield at ||2 mu e_dev|| = fy
    del mu


def test_j2_2d_elastic_branch_is_hooke():
    E, nu = 200.0, 0.3
    m = _mat(_ffi.MAT_ELASTOPLASTIC, E=E, nu=nu, density=1.0, fy=1e9, k1=0.0)
    eps = np.array([[1e-4, 3e-5], [3e-5, -2e-5]])
    sig, st = pyoracle.stress(2, m, eps.T.reshape(-1))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    kappa = lam + mu  # Material.jl:161
    tr = eps[0, 0] + eps[1, 1]
    dev = eps - 0.5 * tr * np.eye(2)
    exp = kappa * tr * np.eye(2) + 2 * mu * dev
    assert np.allclose(sig.reshape(2, 2).T, exp, rtol=1e-13, atol=1e-18)
    assert st[0] == 0.0


def test_johnson_cook_radial_return_lands_on_yield_surface():
    E, nu = 200.0, 0.3
    A, B = 0.3, 1.0
    m = _mat(_ffi.MAT_JOHNSONCOOK, E=E, nu=nu, density=1.0, A=A, B=B, C=0.0, n=1.0, epsdot0=1.0, cellsize=1.0)
    deps = np.zeros((3, 3))
    deps[0, 1] = deps[1, 0] = 5e-3
    sig, st = pyoracle.stress(3, m, deps.T.reshape(-1), deps=deps.T.reshape(-1), dt=1e-3)
    S = sig.reshape(3, 3).T
    s = S - np.trace(S) / 3 * np.eye(3)
    J2 = math.sqrt(1.5) * np.linalg.norm(s)
    assert st[0] > 0.0
    assert abs(J2 - A) < 1e-8  # first plastic step: yield stress A (eff strain < 1e-10 before)


# ---------------------------------------------------------------- whole step
def test_p2g_conserves_mass_and_momentum():
    sc = scenes.block3d(nodes=(10, 10, 10), steps=1)
    eng = OracleEngine(sc.params)
    sc.load(eng)
    eng.step(1, 0.0)
    g = eng.get_grid()
    s = sc.solids[0]
    assert math.isclose(g["mass"].sum(), s["mass"].sum(), rel_tol=1e-13)
    # momentum0 is zeroed on fixed nodes by the grid update: check an
    # unconstrained copy (no Dirichlet mask)
    sc.fixed = None
    eng2 = OracleEngine(sc.params)
    sc.load(eng2)
    eng2.step(1, 0.0)
    g2 = eng2.get_grid()
    assert np.allclose(g2["momentum0"].sum(0), (s["mass"][:, None] * s["vel"]).sum(0), rtol=1e-12)
    # internal forces of a stress field sum to zero (sum_I grad N_I = 0)
    assert np.allclose(g2["force"].sum(0), 0.0, atol=1e-9 * np.abs(g2["force"]).max())


def test_single_particle_p2g_brute_force_2d():
    """One particle: m_I = N_I m, p_I = N_I m v, f_I = -V sigma.gradN_I + m b N_I."""
    p = make_params(2, _ffi.BASIS_QUAD_BSPLINE, _ffi.ALG_USL, 1e-3, (0.0, 0.0), (1.0, 1.0), (11, 11),
                    tolerance=1e-12, body=(0.3, -9.81))
    eng = OracleEngine(p)
    x = np.array([[0.437, 0.552]])
    sig = np.array([[2.0, 0.5, 0.5, -1.0]])
    eng.add_solid(scenes.elastic(1.0, 0.3, 1.0), x, np.array([[0.2, -0.1]]), np.array([2.0]), np.array([0.01]),
                  stress=sig)
    near, N, dN = eng.support(0, 0, True)
    eng.step(1, 0.0)
    g = eng.get_grid()
    S = np.array([[2.0, 0.5], [0.5, -1.0]])
    for k, nid in enumerate(near):
        assert math.isclose(g["mass"][nid], N[k] * 2.0, rel_tol=1e-15)
        assert np.allclose(g["momentum0"][nid], N[k] * 2.0 * np.array([0.2, -0.1]), rtol=1e-15)
        f = -0.01 * S @ dN[k] + 2.0 * np.array([0.3, -9.81]) * N[k]
        assert np.allclose(g["force"][nid], f, rtol=1e-13, atol=1e-16)


def test_uniform_translation_moves_particles_rigidly():
    """Uniform velocity and zero stress: x += v dt, F stays I, stress 0."""
    for dim, basis in ((2, _ffi.BASIS_QUAD_BSPLINE), (3, _ffi.BASIS_QUAD_BSPLINE), (2, _ffi.BASIS_LINEAR)):
        sc = (scenes.block2d if dim == 2 else scenes.block3d)(basis=basis, steps=1)
        sc.fixed = None
        s = sc.solids[0]
        v = np.array([0.7, -0.3, 0.2][:dim])
        s["vel"][:] = v
        s["stress"][:] = 0.0
        sc.params.body[0] = sc.params.body[1] = 0.0
        eng = OracleEngine(sc.params)
        sc.load(eng)
        eng.step(1, 0.0)
        r = eng.get_solid(0)
        assert np.allclose(r["pos"], s["pos"] + v * sc.params.dt, rtol=0, atol=1e-14)
        assert np.allclose(r["vel"], v, atol=1e-12)
        assert np.allclose(r["F"], np.eye(dim).reshape(-1), atol=1e-12)


def test_two_disks_energy_exchange():
    """Sulsky's two-disk collision: KE(0) = 1/2 m v^2 N exactly; the disks
    meet, strain energy takes up most of the kinetic energy, and they
    separate again."""
    g = np.load(os.path.join(GOLD, "two_disks_energy.npz"))["rec"]
    t, ke, se = g[:, 0], g[:, 1], g[:, 2]
    assert len(t) == 350
    assert math.isclose(ke[0], 0.5 * 0.625 * (0.1 ** 2 + 0.1 ** 2) * 416, rel_tol=1e-6)
    assert ke.min() < 0.3 * ke[0]
    assert se.max() > 0.5 * ke[0]
    assert ke[-1] > 0.8 * ke[0]
    # the fixture is reproduced by the restatement as built
    sc = scenes.two_disks()
    eng = OracleEngine(sc.params)
    sc.load(eng)
    tt = 0.0
    tt, _ = eng.step(1, tt)
    assert np.allclose(eng.energies(), (ke[0], se[0]), rtol=1e-15, atol=0)


@pytest.mark.parametrize("material", ["j2", "jc_rate"])
def test_openmp_step_agrees_with_the_serial_restatement(material):
    """The many-core CPU baseline (3D MUSL, particles shared among OpenMP
    threads, scatters coloured by stencil-origin blocks) sums each node block
    by block instead of in particle order: equal to the serial step to
    rounding, on 1 and on 4 threads."""
    from oracle.pyoracle import OracleEngine
    sc = scenes.block3d(nodes=(13, 13, 13), material=material, steps=4)
    a = OracleEngine(sc.params)
    sc.load(a)
    a.step(4, 0.0)
    for threads in (1, 4):
        b = OracleEngine(sc.params)
        sc.load(b)
        b.step_parallel(4, 0.0, threads)
        assert a.flags() == b.flags()
        x, y = a.get_solid(0), b.get_solid(0)
        for k in ("pos", "vel", "F", "strain", "stress", "volume"):
            assert np.max(np.abs(x[k] - y[k])) <= 1e-12 * np.max(np.abs(x[k])), (threads, k)
        ga, gb = a.get_grid(), b.get_grid()
        for k in ("mass", "momentum", "momentum2"):
            assert np.max(np.abs(ga[k] - gb[k])) <= 1e-12 * np.max(np.abs(ga[k])), (threads, k)
