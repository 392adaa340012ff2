"""Pins of the oracle's explicit initial guess (CPU only).

  * the symplectic compositions equal the product of their forward-Euler
    substep matrices (assembled on a tiny grid), and their coefficients are
    the published fourth-order Forest-Ruth / Yoshida values;
  * RK4 (1/6 and 3/8 rules) is fourth order against the exact propagator
    exp(dt A) of df/dt = A f = -{f, H} (halving dt divides the one-step error
    by 32); the symplectic splittings, whose substeps are forward-Euler steps,
    make an O(dt^2) one-step error (their coefficients sum to one);
  * a call with niter substeps moves by dt / niter (the symplectic substep
    matrix), and an RK4 step of size s = dt / niter is exactly the degree-4
    Taylor polynomial of exp(s A) (the linear tendency makes every four-stage
    fourth-order method equal to it);
  * every scheme conserves sum f (sum {f, h} = 0 for fields vanishing near
    the v-boundaries);
  * driver: the guessed Picard run converges to the no-guess solution with no
    more Picard iterations, and a guess that raises the residual is undone.
"""

import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import vlasov_oracle as O

KIND = 'arakawaj4'


def _grid(nx=8, nv=16, ht=0.1, vmax=4.0):
    return O.Grid(nx, nv, 2 * np.pi, vmax, ht=ht)


def _fields(g, seed=0, pad=4):
    rng = np.random.default_rng(seed)
    X, V = np.meshgrid(g.x, g.v, indexing='ij')
    f = (1 + 0.3 * np.cos(X)) * np.exp(-0.5 * V ** 2) + 0.05 * rng.random((g.nx, g.nv))
    f[:, :pad] = f[:, -pad:] = 0.0
    h0 = O.kinetic_hamiltonian(g)
    h1 = 0.2 * np.cos(g.x + 0.3)
    h2 = 0.05 * np.sin(2 * g.x)
    return f, h0, h1, h2


def _matrix(g, h):
    """Dense A with A f = -{f, h} (f flattened in natural order)."""
    N = g.nx * g.nv
    A = np.zeros((N, N))
    for k in range(N):
        e = np.zeros(N)
        e[k] = 1.0
        A[:, k] = O.tendency(KIND, e.reshape(g.nv, g.nx).T, h, g).T.reshape(N)
    return A


def _vec(f):
    return f.T.reshape(-1)


def test_symplectic4_coefficients_are_forest_ruth():
    (c1, c2, c3, c4), (d1, d2, d3) = O.symplectic4_coefficients()
    np.testing.assert_allclose([c1, c2, c3, c4], [0.675603595979829, -0.175603595979829,
                                                  -0.175603595979829, 0.675603595979829], rtol=1e-14)
    np.testing.assert_allclose([d1, d2, d3], [1.351207191959658, -1.702414383919315, 1.351207191959658],
                               rtol=1e-14)


@pytest.mark.parametrize("scheme", ['symplectic2', 'symplectic4'])
def test_symplectic_is_product_of_euler_substeps(scheme):
    g = _grid()
    f, h0, h1, h2 = _fields(g)
    H0 = np.broadcast_to(h0[None, :], (g.nx, g.nv))
    H12 = np.broadcast_to((h1 + h2)[:, None], (g.nx, g.nv))
    Ak, Ap = _matrix(g, H0), _matrix(g, H12)
    I = np.eye(Ak.shape[0])
    dt = g.ht
    if scheme == 'symplectic2':
        seq = [(Ak, 0.5), (Ap, 1.0), (Ak, 0.5)]
        got = O.symplectic_kinetic(KIND, f, H0, 0.5, g)
        got = O.symplectic_potential(KIND, got, H12, 1.0, g)
        got = O.symplectic_kinetic(KIND, got, H0, 0.5, g)
    else:
        c, d = O.symplectic4_coefficients()
        seq = [(Ak, c[0]), (Ap, d[0]), (Ak, c[1]), (Ap, d[1]), (Ak, c[2]), (Ap, d[2]), (Ak, c[3])]
        got = f
        for q in range(3):
            got = O.symplectic_kinetic(KIND, got, H0, c[q], g)
            got = O.symplectic_potential(KIND, got, H12, d[q], g)
        got = O.symplectic_kinetic(KIND, got, H0, c[3], g)
    M = I
    for A, w in seq:
        M = (I + w * dt * A) @ M
    np.testing.assert_allclose(_vec(got), M @ _vec(f), rtol=0, atol=1e-13 * np.abs(f).max())


@pytest.mark.parametrize("niter", [1, 2, 3])
def test_symplectic_substep_divides_the_step(niter):
    """One call with niter substeps moves by factor * dt / niter:
    x -> (I + factor dt / niter A) x, for the kinetic and potential parts."""
    g = _grid()
    f, h0, h1, h2 = _fields(g)
    H0 = np.broadcast_to(h0[None, :], (g.nx, g.nv))
    H12 = np.broadcast_to((h1 + h2)[:, None], (g.nx, g.nv))
    I = np.eye(g.nx * g.nv)
    for step, H in ((O.symplectic_kinetic, H0), (O.symplectic_potential, H12)):
        got = step(KIND, f, H, 0.7, g, niter)
        ref = (I + 0.7 * g.ht / niter * _matrix(g, H)) @ _vec(f)
        np.testing.assert_allclose(_vec(got), ref, rtol=0, atol=1e-13 * np.abs(f).max())


@pytest.mark.parametrize("niter", [1, 3])
@pytest.mark.parametrize("scheme", ['rk4_16', 'rk4_18'])
def test_rk4_is_the_fourth_order_taylor_polynomial(scheme, niter):
    """On df/dt = A f (H fixed, so the tendency is linear) every four-stage
    fourth-order Runge-Kutta step of size s = dt / niter is exactly
    sum_{k<=4} (s A)^k / k!: a wrong stage coefficient or weight changes the
    polynomial."""
    g = _grid()
    f, h0, h1, h2 = _fields(g)
    H = (np.broadcast_to(h0[None, :], (g.nx, g.nv)) + h1[:, None]) + h2[:, None]
    sA = g.ht / niter * _matrix(g, H)
    x, term = _vec(f), _vec(f)
    ref = x.copy()
    for k in range(1, 5):
        term = sA @ term / k
        ref = ref + term
    got = (O.rk4_16 if scheme == 'rk4_16' else O.rk4_18)(KIND, f, H, g, niter)
    np.testing.assert_allclose(_vec(got), ref, rtol=0, atol=1e-13 * np.abs(f).max())


def _one_step_errors(scheme, dts):
    errs = []
    for dt in dts:
        g = _grid(ht=dt)
        f, h0, h1, h2 = _fields(g)
        H0 = np.broadcast_to(h0[None, :], (g.nx, g.nv))
        H12 = np.broadcast_to((h1 + h2)[:, None], (g.nx, g.nv))
        H = (H0 + h1[:, None]) + h2[:, None]
        exact = expm(dt * _matrix(g, H)) @ _vec(f)
        if scheme == 'rk4_16':
            got = O.rk4_16(KIND, f, H, g)
        elif scheme == 'rk4_18':
            got = O.rk4_18(KIND, f, H, g)
        elif scheme == 'symplectic2':
            got = O.symplectic_kinetic(KIND, f, H0, 0.5, g)
            got = O.symplectic_potential(KIND, got, H12, 1.0, g)
            got = O.symplectic_kinetic(KIND, got, H0, 0.5, g)
        else:
            c, d = O.symplectic4_coefficients()
            got = f
            for q in range(3):
                got = O.symplectic_kinetic(KIND, got, H0, c[q], g)
                got = O.symplectic_potential(KIND, got, H12, d[q], g)
            got = O.symplectic_kinetic(KIND, got, H0, c[3], g)
        errs.append(np.abs(_vec(got) - exact).max())
    return errs


@pytest.mark.parametrize("scheme,ratio", [('rk4_16', 32.0), ('rk4_18', 32.0), ('symplectic2', 4.0),
                                          ('symplectic4', 4.0)])
def test_one_step_error_order(scheme, ratio):
    e = _one_step_errors(scheme, [0.02, 0.01, 0.005])
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 0.8 * ratio < r1 < 1.25 * ratio and 0.8 * ratio < r2 < 1.25 * ratio, (scheme, e, r1, r2)


@pytest.mark.parametrize("scheme", ['rk4_16', 'rk4_18', 'symplectic2', 'symplectic4'])
def test_mass_is_conserved(scheme):
    """Each stage widens the support by the stencil's two rows, so f vanishes
    within 14 rows of the v-boundaries (no stage then reaches the unbracketed rows)."""
    g = _grid(16, 64, ht=0.05, vmax=6.0)
    f, h0, h1, h2 = _fields(g, 3, pad=14)
    H0 = np.broadcast_to(h0[None, :], (g.nx, g.nv))
    H = (H0 + h1[:, None]) + h2[:, None]
    if scheme.startswith('rk'):
        got = (O.rk4_16 if scheme == 'rk4_16' else O.rk4_18)(KIND, f, H, g)
    else:
        got = O.symplectic_potential(KIND, f, np.broadcast_to((h1 + h2)[:, None], f.shape), 1.0, g)
        got = O.symplectic_kinetic(KIND, got, H0, 0.5, g)
    assert abs(got.sum() - f.sum()) <= 1e-13 * f.sum()


def _landau(guess, niter=1, steps=3):
    g = O.Grid(16, 64, 4 * np.pi, 8.0, ht=0.1)
    prof = lambda x, v: (1 + 0.3 * np.cos(0.5 * x)) * np.exp(-0.5 * (v - 0.5) ** 2) / math.sqrt(2 * np.pi)
    run = O.OracleRun(g, O.initial_distribution(g, prof), ksp_rtol=1e-12, snes_atol=1e-11,
                      initial_guess=guess, initial_iter=niter)
    log = [run.step() for _ in range(steps)]
    return run, log


@pytest.mark.parametrize("guess,niter", [('symplectic2', 1), ('symplectic4', 2), ('rk4', 1), ('rk4_16', 1)])
def test_guessed_run_converges_to_the_same_solution(guess, niter):
    base, lb = _landau(None)
    run, lg = _landau(guess, niter)
    assert np.abs(run.fc - base.fc).max() <= 1e-9 * np.abs(base.fc).max()
    # step 1: the solver's H_ave is still the zero Vec it was created as (no
    # update_previous before the first step), so F(fh) = 0 and any guess
    # raises the residual: it is undone, as in the reference's driver
    prev, ig = lg[0]['guess']
    assert ig > prev
    for eb, eg in zip(lb[1:], lg[1:]):
        prev, ig = eg['guess']
        assert ig < prev                       # the guess is kept: the residual went down
        assert eg['initial'] < eb['initial']   # the Picard loop starts closer
        assert len(eg['picard']) <= len(eb['picard'])


def test_guess_that_raises_the_residual_is_undone():
    g = O.Grid(16, 64, 4 * np.pi, 8.0, ht=0.1)
    prof = lambda x, v: (1 + 0.3 * np.cos(0.5 * x)) * np.exp(-0.5 * v ** 2) / math.sqrt(2 * np.pi)
    run = O.OracleRun(g, O.initial_distribution(g, prof), initial_guess='symplectic2')
    run.fh = run.fc.copy()
    run.h1h, run.h2h = run.h1c.copy(), run.h2c.copy()
    f0 = run.fc.copy()
    # a guess that moves f far from the history state raises the residual
    run._guess_iteration = lambda: setattr(run, 'fc', run.fc + 1.0)
    prev, ig = run.initial_guess()
    assert ig > prev and np.array_equal(run.fc, f0)
