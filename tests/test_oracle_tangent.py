"""Pins of the tangent-linear oracle (CPU).

G(f1, f0) is quadratic in f1 and in f0 (B bilinear, h1 linear), so a
central difference of G is exact up to rounding: it pins A and A0 without
sharing any code path with their assembly.  The field-energy gradient is
pinned the same way (E quadratic in f), and the whole tangent-linear step
against central differences of full Picard timesteps of the oracle run.
"""

import math

import numpy as np
import pytest

from oracle import tangent_linear as TL
from oracle import vlasov_oracle as O

KINDS = ['arakawaj1', 'arakawaj2', 'arakawaj4']


def _pair(nx, nv, seed):
    rng = np.random.default_rng(seed)
    g = O.Grid(nx, nv, 2 * math.pi / 0.3, 8.0, ht=0.05)
    f0 = rng.random((nx, nv))
    f0[:, :2] = 0
    f0[:, -2:] = 0
    f1 = f0 + 0.05 * rng.standard_normal((nx, nv))
    f1[:, :2] = 0
    f1[:, -2:] = 0
    return g, f0, f1, rng


@pytest.mark.parametrize("kind", KINDS)
def test_linearisation_is_the_central_difference(kind):
    g, f0, f1, rng = _pair(16, 24, 1)
    h0 = O.kinetic_hamiltonian(g)
    u = rng.standard_normal(f0.shape)
    eps = 1e-3
    for which, shift in (('A', lambda e: (f1 + e * u, f0)), ('A0', lambda e: (f1, f0 + e * u))):
        gp = TL.step_residual(kind, *shift(eps), h0, g, -1.0)
        gm = TL.step_residual(kind, *shift(-eps), h0, g, -1.0)
        fd = (gp - gm) / (2 * eps)
        lin = TL.linear_apply(kind, u, f0, f1, h0, g, -1.0, which)
        assert np.abs(fd - lin).max() <= 1e-8 * np.abs(lin).max(), which
    # a dropped coupling term would fail this by O(1)
    H = O.h_average(h0, TL.h1_of(f1, g, -1.0), TL.h1_of(f0, g, -1.0), np.zeros(g.nx), np.zeros(g.nx))
    assert np.abs(lin - O.jacobian(kind, u, H, g)).max() > 1e-3 * np.abs(lin).max()


def test_dense_matches_matrix_free():
    g, f0, f1, rng = _pair(8, 16, 2)
    h0 = O.kinetic_hamiltonian(g)
    A = TL.dense('arakawaj4', f0, f1, h0, g, -1.0, 'A')
    u = rng.standard_normal(f0.shape)
    assert np.allclose(A @ u.ravel(), TL.linear_apply('arakawaj4', u, f0, f1, h0, g, -1.0, 'A').ravel(),
                       rtol=0, atol=1e-12 * np.abs(A).max() * np.abs(u).max())


def test_field_energy_gradient():
    g, f0, f1, rng = _pair(16, 24, 3)
    d = rng.standard_normal(f0.shape)
    eps = 1e-3
    fd = (TL.field_energy(f1 + eps * d, g, -1.0) - TL.field_energy(f1 - eps * d, g, -1.0)) / (2 * eps)
    an = (TL.field_energy_gradient(f1, g, -1.0) * d).sum()
    assert abs(fd - an) <= 1e-9 * abs(an)


def test_tangent_linear_matches_picard_differences():
    """d_K of the tangent-linear model = central difference of K converged
    timesteps (tight Picard / GMRES tolerances so the forward map is smooth
    to rounding)."""
    nx, nv, K = 8, 16, 2
    g = O.Grid(nx, nv, 2 * math.pi / 0.3, 8.0, ht=0.05)
    prof = lambda x, v: (1 + 0.3 * np.cos(0.3 * x)) * np.exp(-0.5 * v ** 2) / math.sqrt(2 * math.pi)
    base = O.normalise(O.initial_distribution(g, prof), g)
    rng = np.random.default_rng(4)
    d0 = rng.standard_normal((nx, nv))
    d0[:, :2] = 0
    d0[:, -2:] = 0

    def run(f0):
        r = O.OracleRun(g, f0, bracket='arakawaj4', ksp_rtol=1e-14, snes_atol=1e-14, snes_max_iter=30)
        r.fc = np.array(f0)            # no re-normalisation of the perturbed state
        r.calculate_moments()
        states = [r.fc.copy()]
        for _ in range(K):
            r.step()
            states.append(r.fc.copy())
        return states

    states = run(base)
    h0 = O.kinetic_hamiltonian(g)
    dK = TL.tangent_linear('arakawaj4', states, d0, h0, g, -1.0)
    eps = 1e-6
    fp, fm = run(base + eps * d0)[-1], run(base - eps * d0)[-1]
    fd = (fp - fm) / (2 * eps)
    err = np.abs(fd - dK).max() / np.abs(dK).max()
    assert err <= 1e-6, err


def test_field_energy_has_the_reference_charge_factor():
    """potential.pyx:48-59: E = 0.5 * charge * sum (phi_{i+1} - phi_i)^2 / hx;
    with q = +1 the sum itself, with q = -1 its negative (phi itself flips
    sign with q, so the squares do not)."""
    g, f0, f1, rng = _pair(16, 24, 5)
    phi = O.poisson_solve(O.poisson_rhs(O.density(f1, g), 1.0, g.nx), g.hx, 'cfd4')
    s = 0.0
    for i in range(g.nx):
        s += (phi[(i + 1) % g.nx] - phi[i]) ** 2
    assert abs(TL.field_energy(f1, g, 1.0) - 0.5 * s / g.hx) <= 1e-13 * abs(s / g.hx)
    assert abs(TL.field_energy(f1, g, -1.0) + 0.5 * s / g.hx) <= 1e-13 * abs(s / g.hx)


def test_matrix_free_tangent_linear_matches_dense():
    """The matrix-free model (SciPy GMRES on linear_apply, kinetic P^-1)
    propagates a perturbation through 3 steps like the dense one."""
    nx, nv = 16, 32
    g = O.Grid(nx, nv, 2 * math.pi / 0.3, 8.0, ht=0.05)
    prof = lambda x, v: (1 + 0.3 * np.cos(0.3 * x)) * np.exp(-0.5 * v ** 2) / math.sqrt(2 * math.pi)
    r = O.OracleRun(g, O.normalise(O.initial_distribution(g, prof), g), bracket='arakawaj4', ksp_rtol=1e-14,
                    snes_atol=1e-14, snes_max_iter=30)
    states = [r.fc.copy()]
    for _ in range(3):
        r.step()
        states.append(r.fc.copy())
    h0 = O.kinetic_hamiltonian(g)
    rng = np.random.default_rng(8)
    d0 = rng.standard_normal((nx, nv))
    dense = TL.tangent_linear('arakawaj4', states, d0, h0, g, -1.0)
    mf = TL.tangent_linear_mf('arakawaj4', states, d0, h0, g, -1.0)
    err = np.abs(mf - dense).max() / np.abs(dense).max()
    assert err <= 1e-10, err
