"""Pins of the oracle's preconditioner, Poisson solve, moments, invariants,
GMRES and Picard driver (CPU only)."""

import math

import numpy as np
import pytest

from oracle import stencil_tables as ST
from oracle import vlasov_oracle as O


# ---------------------------------------------------------------- preconditioner
def test_eigenvalues_are_centered_difference_symbol():
    """ifftshift(exp(2 pi i k (n-1)/n) - exp(2 pi i k/n)) = 2 i sin(2 pi k/n)
    for even n (the symbol of f_{i+1} - f_{i-1}); for odd n ifftshift
    rotates by (n-1)/2, which gives 2 i sin(pi (2k - 1)/n)."""
    for n in (8, 10, 64, 100, 1024):
        e = O.kinetic_eigenvalues(n)
        k = np.arange(n)
        np.testing.assert_allclose(e, 2j * np.sin(2 * np.pi * k / n), rtol=0, atol=1e-11)
    for n in (9, 51, 201):
        e = O.kinetic_eigenvalues(n)
        k = np.arange(n)
        np.testing.assert_allclose(e, 2j * np.sin(np.pi * (2 * k - 1) / n), rtol=0, atol=1e-11)


def test_lapack_and_scipy_statements_agree():
    g = O.Grid(32, 64, 4 * np.pi, 8.0, ht=0.05)
    a, b = O.Preconditioner(g, 'lapack'), O.Preconditioner(g, 'scipy')
    x = np.random.default_rng(0).standard_normal((g.nx, g.nv))
    ya, yb = a.apply(x), b.apply(x)
    assert np.abs(ya - yb).max() <= 1e-13 * np.abs(ya).max()
    assert a.pivots_are_identity()


def test_pivots_identity_at_benchmark_sizes():
    for (nx, nv, L, vmax, ht) in [(64, 128, 4 * np.pi, 10., 0.1), (256, 512, 4 * np.pi, 8., 0.05)]:
        assert O.Preconditioner(O.Grid(nx, nv, L, vmax, ht=ht)).pivots_are_identity()


def test_odd_length_preconditioner_is_a_modewise_dense_solve():
    """Odd nx: the reference's eigenvalues make T_0 complex (2 i sin(-pi/nx)),
    and the c2r transform drops Im of mode 0, so P^-1 is not the inverse of
    the forward composition; it is still irfft(T_k^-1 rfft(x)) mode by mode,
    checked here against dense complex solves of the bands."""
    for nx, nv in ((9, 10), (51, 12)):
        g = O.Grid(nx, nv, 2 * np.pi, 4.0, ht=0.1)
        pc = O.Preconditioner(g)
        x = np.random.default_rng(nx).standard_normal((nx, nv))
        X = np.fft.rfft(x, axis=0)
        for k in range(nx // 2 + 1):
            dm, d, dp = O.kinetic_bands(g, O.kinetic_eigenvalues(nx)[k])
            X[k] = np.linalg.solve(np.diag(d) + np.diag(dp[:-1], 1) + np.diag(dm[1:], -1), X[k])
        ref = np.fft.irfft(X, n=nx, axis=0)
        assert np.abs(pc.apply(x) - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("nx,nv", [(8, 16), (10, 12)])
def test_p_times_pinv_is_identity(nx, nv):
    g = O.Grid(nx, nv, 2 * np.pi, 4.0, ht=0.1)
    pc = O.Preconditioner(g)
    P = pc.dense_matrix()
    x = np.random.default_rng(1).standard_normal((g.nx, g.nv))
    y = pc.apply(x)
    Py = (P @ y.T.reshape(-1)).reshape((g.nv, g.nx)).T
    assert np.abs(Py - x).max() <= 1e-12 * np.abs(x).max()


@pytest.mark.parametrize("nx", [8, 10])
def test_preconditioner_is_half_j1_of_h0_plus_identity_over_dt(nx):
    """On interior rows P u = u/dt + 1/2 B_J1(u, h0): the bands of
    KineticFast.formBandedPreconditionerMatrix are the Fourier symbol of the
    J1 bracket with the kinetic Hamiltonian (checked through the stencil
    tables, not the oracle's bracket), for even nx."""
    g = O.Grid(nx, 16, 2 * np.pi, 4.0, ht=0.1)
    pc = O.Preconditioner(g)
    P = pc.dense_matrix()
    H0 = np.broadcast_to(O.kinetic_hamiltonian(g)[None, :], (g.nx, g.nv)).copy()
    M = ST.bracket_matrix('arakawaj1', H0, g, factor=0.5) + np.eye(g.nx * g.nv) / g.ht
    interior = np.zeros((g.nv, g.nx), bool)
    interior[2:-2, :] = True
    rows = interior.reshape(-1)
    assert np.abs(P[rows] - M[rows]).max() <= 1e-12 * np.abs(M).max()


# ---------------------------------------------------------------- Poisson
@pytest.mark.parametrize("nx", [16, 9, 20])
@pytest.mark.parametrize("laplace", ['cfd4', 'cfd2'])
def test_poisson_solve_vs_dense_lstsq(laplace, nx):
    hx = 0.37
    A = np.zeros((nx, nx))
    for c in range(nx):
        e = np.zeros(nx)
        e[c] = 1
        A[:, c] = O.poisson_matrix_mult(e, 1 / hx ** 2, laplace)
    # stencil of PETScPoissonCFD4.formMat / CFD2.formMat
    row = A[5]
    if laplace == 'cfd4':
        np.testing.assert_allclose(row[[3, 4, 5, 6, 7]], np.array([5, -32, 54, -32, 5]) / (12 * hx ** 2))
    else:
        np.testing.assert_allclose(row[[4, 5, 6]], np.array([-1, 2, -1]) / hx ** 2)
    n = 1 + 0.3 * np.random.default_rng(2).standard_normal(nx)
    b = O.poisson_rhs(n, -1.0, nx)
    phi = O.poisson_solve(b, hx, laplace)
    ref = np.linalg.lstsq(A, b, rcond=None)[0]
    np.testing.assert_allclose(phi - phi.mean(), ref - ref.mean(), rtol=0, atol=1e-12 * np.abs(ref).max())
    r = O.poisson_function_mult(phi, n, -1.0, 1 / hx ** 2, laplace)
    assert np.linalg.norm(r) <= 1e-12 * np.linalg.norm(n)


def test_poisson_eigenvalues():
    nx, hx = 32, 0.2
    lam = O.poisson_eigenvalues(nx, hx, 'cfd4')
    for k in (1, 5, 16):
        e = np.cos(2 * np.pi * k * np.arange(nx) / nx)
        np.testing.assert_allclose(O.poisson_matrix_mult(e, 1 / hx ** 2, 'cfd4'), lam[k] * e, atol=1e-10)


# ---------------------------------------------------------------- moments
def test_density_and_normalisation_of_maxwellian():
    g = O.Grid(16, 257, 4 * np.pi, 10.0)
    f = O.initial_distribution(g, lambda x, v: O.maxwellian(v) * (1 + 0.1 * np.cos(0.5 * x)))
    n = O.density(f, g)
    np.testing.assert_allclose(n, 1 + 0.1 * np.cos(0.5 * g.x), rtol=1e-12)
    f2 = O.normalise(f, g)
    assert abs(f2.sum() * g.hv / g.nx - 1.0) <= 1e-14
    assert np.all(f2[:, :2] == 0) and np.all(f2[:, -2:] == 0)


def test_velocity_and_energy_density_of_drifting_maxwellian():
    """u = n u0 and e = n (1 + u0^2) for f = n(x) M(v - u0) (the rectangle rule
    is spectrally accurate for a Gaussian that vanishes at the v-boundary); a
    wrong power of v or a missing hv fails by O(1)."""
    g = O.Grid(16, 257, 4 * np.pi, 10.0)
    u0 = 0.7
    nx_prof = 1 + 0.1 * np.cos(0.5 * g.x)
    f = O.initial_distribution(g, lambda x, v: O.maxwellian(v, vd=u0) * (1 + 0.1 * np.cos(0.5 * x)))
    np.testing.assert_allclose(O.velocity_density(f, g), nx_prof * u0, rtol=1e-12)
    np.testing.assert_allclose(O.energy_density(f, g), nx_prof * (1 + u0 ** 2), rtol=1e-12)
    # symmetric in v: u vanishes to rounding
    fs = O.initial_distribution(g, lambda x, v: O.maxwellian(v) * (1 + 0.1 * np.cos(0.5 * x)))
    assert np.abs(O.velocity_density(fs, g)).max() <= 1e-14


def test_invariants_closed_form():
    g = O.Grid(32, 401, 2 * np.pi, 10.0)
    f = np.broadcast_to(O.maxwellian(g.v)[None, :], (g.nx, g.nv)).copy()
    h0 = O.kinetic_hamiltonian(g)
    z = np.zeros(g.nx)
    inv = O.invariants(f, h0, z, z, z, g, -1.0)
    L = g.L
    assert abs(inv['mass'] - L) <= 1e-10 * L
    assert abs(inv['l2'] - L / (2 * math.sqrt(np.pi))) <= 1e-10 * L
    assert abs(inv['energy'] - 0.5 * L) <= 1e-10 * L
    phi = np.sin(2 * np.pi * g.x / L)
    inv2 = O.invariants(f, h0, z, z, phi, g, -1.0)
    # 1/2 q sum (dphi)^2 / hx  ~  q/2 * int phi'^2 dx = q/2 * (2pi/L)^2 * L/2
    exact = -0.5 * (2 * np.pi / L) ** 2 * L / 2
    assert abs(inv2['efield'] - exact) <= 1e-2 * abs(exact)


# ---------------------------------------------------------------- GMRES
def _system(n, seed, cond=10.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.linspace(1, cond, n)) @ Q.T + 0.3 * rng.standard_normal((n, n)) / math.sqrt(n)
    return A, rng.standard_normal(n)


@pytest.mark.parametrize("restart", [50, 10, 3])
def test_gmres_vs_dense_solve(restart):
    A, b = _system(40, 3)
    hist = []
    x, its, reason = O.gmres(lambda v: A @ v, b, rtol=1e-12, atol=1e-50, max_it=5000, restart=restart,
                             history=hist)
    assert reason == O.KSP_CONVERGED_RTOL
    ref = np.linalg.solve(A, b)
    assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max()
    assert np.linalg.norm(b - A @ x) <= 1e-11 * np.linalg.norm(b)
    if restart >= 40:
        assert its <= 41
    # residual estimates never increase (GMRES minimises over a growing space;
    # a restart recomputes b - Ax, equal up to rounding of size eps ||b||)
    assert all(hist[i + 1] <= hist[i] * (1 + 1e-12) + 1e-13 * hist[0] for i in range(len(hist) - 1))


def test_gmres_full_restart_equivalence():
    A, b = _system(25, 4)
    x1, i1, _ = O.gmres(lambda v: A @ v, b, rtol=1e-10, restart=30)
    x2, i2, _ = O.gmres(lambda v: A @ v, b, rtol=1e-10, restart=100)
    assert i1 == i2 and np.abs(x1 - x2).max() <= 1e-12 * np.abs(x1).max()


def test_gmres_edge_cases():
    A, b = _system(10, 5)
    x, its, reason = O.gmres(lambda v: A @ v, np.zeros(10))
    assert reason == O.KSP_CONVERGED_ATOL and its == 0 and not x.any()
    x, its, reason = O.gmres(lambda v: A @ v, b, rtol=1e-14, max_it=3, restart=30)
    assert reason == O.KSP_DIVERGED_ITS and its == 3
    x, its, reason = O.gmres(lambda v: v, b, rtol=1e-10)     # identity: happy breakdown / one step
    assert reason > 0 and its == 1 and np.abs(x - b).max() <= 1e-14


# ---------------------------------------------------------------- driver
def _landau_run(**kw):
    g = O.Grid(16, 32, 4 * np.pi, 8.0, ht=0.05)
    prof = lambda x, v: (1 + 0.3 * np.cos(0.5 * x)) * O.maxwellian(v)
    return O.OracleRun(g, O.initial_distribution(g, prof), **kw)


def test_driver_converges_and_conserves():
    run = _landau_run()
    inv0 = O.invariants(run.fc, run.h0, run.h1c, run.h2c, run.phi, run.grid, -1.0)
    for _ in range(3):
        e = run.step()
        assert e['picard'][-1]['residual'] < 1e-10 or len(e['picard']) == 20
    inv1 = O.invariants(run.fc, run.h0, run.h1c, run.h2c, run.phi, run.grid, -1.0)
    # f vanishes near the v-boundaries, so the J4 bracket conserves mass and L2
    assert abs(inv1['mass'] - inv0['mass']) <= 1e-10 * inv0['mass']
    assert abs(inv1['l2'] - inv0['l2']) <= 1e-8 * inv0['l2']


def test_driver_max_iter_branch():
    run = _landau_run(snes_max_iter=1)
    e = run.step()
    assert len(e['picard']) == 1


def test_driver_revert_branch():
    run = _landau_run(snes_atol=0.0, snes_max_iter=50)
    e = run.step()
    # with snes_atol = 0 the loop only ends when the residual stops decreasing
    # (rounding floor) or after max_iter; a rise reverts f to the last iterate
    assert len(e['picard']) < 50
    last = e['picard'][-1]['residual']
    assert last >= e['picard'][-2]['residual']


def _prescribed_run(residuals, **kw):
    """The oracle's Picard loop with the solve replaced by f += 1 and the
    residual sequence prescribed (the pre-loop residual first)."""
    run = _landau_run(**kw)
    seq = iter(residuals)
    run.calculate_residual = lambda: next(seq)
    run.calculate_moments = lambda: None
    run.update_previous = lambda: None

    def solve():
        run.fc = run.fc + 1.0
        return 7, 2
    run.snes_ksponly = solve
    return run


# (residuals, snes_atol, snes_max_iter) -> (Picard iterations, iterates kept)
# by run_nonlinear_matrixfree_split_mp.py:158-163: stop when the residual
# rose after the first iteration, fell below snes_atol or hit the cap; a rise
# at the stop (the first iteration included) undoes the last iterate
PRESCRIBED = [
    (([1.0, 1e-3, 1e-7, 5e-11], 1e-10, 20), (3, 3)),      # converges
    (([1.0, 1e-3, 2e-3, 1e-4], 1e-10, 20), (2, 1)),       # rise: reverted
    (([1.0, 2.0, 1e-11], 1e-10, 20), (2, 2)),             # a first-iteration rise does not stop
    (([2e-11, 5e-11, 1e-12], 1e-10, 20), (1, 0)),         # ... but is undone when below snes_atol
    (([1.0, 2.0, 1e-3], 1e-10, 1), (1, 0)),               # ... or at the cap
    (([1.0] + [0.5 ** k for k in range(1, 30)], 1e-10, 5), (5, 5)),   # the cap
    (([1.0, 1e-10, 1e-12], 1e-10, 20), (2, 2)),           # snes_atol itself does not stop
    (([1.0, 1e-3, 1e-10, 1e-12], 1e-10, 20), (3, 3)),
    (([1.0, 3.0, 0.5, 1e-12], 1e-10, 20), (3, 3)),        # a first-iteration rise, then convergence
]


@pytest.mark.parametrize("case,expect", PRESCRIBED)
def test_driver_stop_and_revert_rule(case, expect):
    residuals, atol, cap = case
    run = _prescribed_run(residuals, snes_atol=atol, snes_max_iter=cap)
    f0 = run.fc.copy()
    e = run.step()
    assert (len(e['picard']), int(round((run.fc - f0).max()))) == expect


def test_driver_abort_branch(monkeypatch):
    run = _landau_run()
    monkeypatch.setattr(run, 'calculate_residual', lambda: 100.0)
    with pytest.raises(RuntimeError):
        run.step()


def _potential_case(nx=8, nv=16, seed=0):
    g = O.Grid(nx, nv, 2 * np.pi, 5.0, ht=0.1)
    phi = np.random.default_rng(seed).standard_normal(nx) * 0.3
    return g, phi


def test_potential_preconditioner_is_minus_half_j1_of_phi_plus_identity_over_dt():
    """Interior rows of P (1 <= i <= nx-2, 2 <= j <= nv-3) equal
    I/dt - 1/2 J1(., phi) from the pinned bracket oracle; the other rows differ
    (dropped periodic corners in x, periodic v in the FFT, masked bracket
    rows).  The sign is the reference's: its bands put +e*(phi differences)
    where the bracket {u, phi} = -u_v phi_x carries a minus."""
    g, phi = _potential_case()
    pc = O.PotentialPreconditioner(g, phi)
    P = pc.dense_matrix()
    nx, nv = g.nx, g.nv
    N = nx * nv
    h = np.broadcast_to(phi[:, None], (nx, nv)).copy()
    J = np.zeros((N, N))
    for col in range(N):
        e = np.zeros(N)
        e[col] = 1.0
        u = e.reshape((nv, nx)).T
        y = g.ht_inv * u
        O.bracket_add('arakawaj1', u, h, y, -0.5, g)
        J[:, col] = y.T.reshape(N)
    d = np.abs(P - J).max(axis=1).reshape(nv, nx)
    interior = np.zeros((nv, nx), dtype=bool)
    interior[2:nv - 2, 1:nx - 1] = True
    assert d[interior].max() <= 1e-14 * np.abs(P).max()
    assert d[~interior].min() > 1e-6        # every other row differs


def test_potential_preconditioner_variants_and_inverse():
    g, phi = _potential_case(16, 32, 1)
    x = np.random.default_rng(2).standard_normal((16, 32))
    a = O.PotentialPreconditioner(g, phi, 'lapack')
    b = O.PotentialPreconditioner(g, phi, 'scipy')
    assert a.pivots_are_identity()
    np.testing.assert_allclose(a.apply(x), b.apply(x), rtol=0, atol=1e-13 * np.abs(a.apply(x)).max())
    g8, phi8 = _potential_case(8, 16, 3)
    c = O.PotentialPreconditioner(g8, phi8)
    x8 = np.random.default_rng(4).standard_normal((8, 16))
    y = c.dense_matrix() @ c.apply(x8).T.reshape(-1)
    assert np.abs(y - x8.T.reshape(-1)).max() <= 1e-13 * np.abs(x8).max()
    # update() rebuilds the bands from a new phi
    c.update(2 * phi8)
    assert np.allclose(c.bands[1][0], O.potential_bands(g8, 2 * phi8, c.eigen[1])[0])


def test_driver_with_potential_preconditioner():
    """The Picard driver with tensor_potential converges to the same step as
    with tensor_kinetic (the preconditioner changes the iterates, not the
    converged solution)."""
    g = O.Grid(16, 32, 4 * np.pi, 8.0, ht=0.05)
    prof = lambda x, v: (1 + 0.5 * np.cos(0.5 * x)) * O.maxwellian(v)
    f0 = O.initial_distribution(g, prof)
    rk = O.OracleRun(g, f0, ksp_rtol=1e-13, snes_atol=1e-13)
    rp = O.OracleRun(g, f0, ksp_rtol=1e-13, snes_atol=1e-13, preconditioner='tensor_potential')
    for _ in range(2):
        rk.step()
        ep = rp.step()
        assert all(p['reason'] > 0 for p in ep['picard'])
    assert np.abs(rk.fc - rp.fc).max() <= 1e-10 * np.abs(rk.fc).max()
