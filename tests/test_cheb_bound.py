"""The facts the 2D MUSL solve rests on, checked on the CPU.

The device solves the blended mass system B x = b, B = 0.1 diag(m) + 0.9 M
(M_IJ = sum_p m_p N_I N_J, m_I = sum_p m_p N_I), by a fixed number of
Chebyshev iterations preconditioned with D = diag(m) on the interval
[0.1, 1] (the Chebyshev kernels in jump_amd/csrc/mpm_kernels.hpp, the
count in solve_musl2d of jump_amd/csrc/jump_api.hip).  Checked here on
random particle layouts with the quadratic B-spline (and the linear hat):
the spectrum of D^-1 B lies in [0.1, 1], the iteration with the library's
coefficients and count reaches the dense solution to the bound, and an
interval that misses the spectrum's low end does not."""
import numpy as np
import pytest


def _qbs(r):
    a = np.abs(r)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def _hat(r):
    return np.maximum(1.0 - np.abs(r), 0.0)


def _system(basis, n=12, npart=400, seed=0):
    rng = np.random.default_rng(seed)
    nn = n * n
    M = np.zeros((nn, nn))
    ii = np.arange(n)
    for _ in range(npart):
        x, y = rng.uniform(0.0, n - 1.0, 2)  # near the edges some nodes fall outside: sum N < 1
        mp = rng.uniform(0.5, 1.5)
        N = np.outer(basis(y - ii), basis(x - ii)).ravel()  # node i + n j
        M += mp * np.outer(N, N)
    m = M.sum(1)  # the lumped mass: M's row sums
    B = 0.1 * np.diag(m) + 0.9 * M
    mass = m > 0.0
    B[~mass, :] = 0.0
    B[~mass, ~mass] = 1.0  # massless rows are the identity
    d = np.where(mass, m, 1.0)
    return B, d, mass, rng


def _iterations(tol=1e-14):
    q = (np.sqrt(10.0) - 1.0) / (np.sqrt(10.0) + 1.0)
    k = 0
    while True:
        k += 1
        if 2.0 * q**k / (1.0 + q ** (2 * k)) <= tol:
            return k


def _chebyshev(B, d, b, K, lo=0.1, hi=1.0):
    theta, delta = (hi + lo) / 2, (hi - lo) / 2
    sigma = theta / delta
    x = np.zeros_like(b)
    r = b.copy()
    dk = r / d / theta
    rho = 1.0 / sigma
    for _ in range(K):
        x = x + dk
        r = r - B @ dk
        rn = 1.0 / (2.0 * sigma - rho)
        dk = rn * rho * dk + (2.0 * rn / delta) * (r / d)
        rho = rn
    return x


def test_the_iteration_count_is_51():
    assert _iterations() == 51


@pytest.mark.parametrize("basis", [_qbs, _hat])
@pytest.mark.parametrize("seed", [0, 1])
def test_spectrum_lies_in_the_interval(basis, seed):
    B, d, mass, _ = _system(basis, seed=seed)
    ev = np.linalg.eigvals(B / d[:, None]).real
    assert ev.min() >= 0.1 * (1 - 1e-12)
    assert ev.max() <= 1.0 * (1 + 1e-12)


@pytest.mark.parametrize("basis", [_qbs, _hat])
def test_chebyshev_reaches_the_dense_solution(basis):
    B, d, mass, rng = _system(basis, seed=3)
    b = np.where(mass, rng.normal(size=len(d)), 0.0)
    xs = np.linalg.solve(B, b)
    x = _chebyshev(B, d, b, _iterations())
    e = x - xs
    assert np.sqrt(e @ B @ e) <= 1e-13 * np.sqrt(xs @ B @ xs)
    # a uniform velocity field is solved exactly up to rounding: B 1 = D 1 on massive rows
    u = np.where(mass, 1.0, 0.0)
    xu = _chebyshev(B, d, B @ (0.3 * u), _iterations())
    assert np.max(np.abs(xu[mass] - 0.3)) <= 1e-14


def test_an_interval_missing_the_low_end_fails():
    # the same iteration on [0.3, 1] leaves the low modes unconverged: the
    # bound's lower end matters
    B, d, mass, rng = _system(_qbs, seed=4)
    b = np.where(mass, rng.normal(size=len(d)), 0.0)
    xs = np.linalg.solve(B, b)
    x = _chebyshev(B, d, b, _iterations(), lo=0.3)
    e = x - xs
    assert np.sqrt(e @ B @ e) > 1e-10 * np.sqrt(xs @ B @ xs)
