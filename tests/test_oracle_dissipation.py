"""Pins of the oracle's dissipation terms (CPU only).

Collision operators (CollisionOperator.pyx:68-250) and the collision factor
(Toolbox.pyx:153-165):
  * the reference's second statement of the operators, the (j-1, j, j+1)
    entries of its assembled jacobian();
  * exact discrete conservation of mass, momentum and energy when the moments
    are those of f itself and f vanishes near the v-boundaries (a dropped 0.5,
    hv in place of hv^2, a wrong collision factor or index breaks it);
  * a drifting Maxwellian is in the null space to second order in hv, and
    the diffusion part alone has sum f C(f) > 0 (the dissipative sign);
  * a = 1/T for a Maxwellian.
Double brackets (DoubleBracket.pyx:67-97):
  * the fully antisymmetric Arakawa trilinear form gives
    sum H D_ffh = nu sum {F,H}^2, sum F D_ffh = 0, sum F D_fhh = nu sum {F,H}^2,
    sum H D_fhh = 0 (a swapped argument flips or breaks them);
  * the Jacobian terms are the exact derivative of the function terms
    (quadratic / linear in f: central differences are exact up to rounding).
"""

import numpy as np
import pytest

from oracle import vlasov_oracle as O

BRACKETS = ['arakawaj1', 'arakawaj2', 'arakawaj4']
COLLS = ['collt', 'colle']


def _grid(nx=16, nv=64, L=2 * np.pi, vmax=8.0, ht=0.1):
    return O.Grid(nx, nv, L, vmax, ht=ht)


def _maxwellians(g, n0, u0, T0):
    """f[i, j] = n0_i / sqrt(2 pi T0_i) exp(-(v_j - u0_i)^2 / (2 T0_i))."""
    v = g.v[None, :]
    return n0[:, None] / np.sqrt(2 * np.pi * T0[:, None]) * np.exp(-(v - u0[:, None]) ** 2 / (2 * T0[:, None]))


def _moments(f, g):
    n, u, e = O.density(f, g), O.velocity_density(f, g), O.energy_density(f, g)
    return n, u, e, O.collision_factor(n, u, e)


def _compact_random(g, rng, pad=4):
    f = 1.0 + rng.random((g.nx, g.nv))
    f[:, :pad] = f[:, -pad:] = 0.0
    return f


def test_collision_factor_is_inverse_temperature():
    g = _grid(8, 400, vmax=12.0)
    x = np.arange(g.nx)
    n0, u0, T0 = 1.0 + 0.2 * np.cos(x), 0.3 * np.sin(x), 0.8 + 0.3 * np.cos(2 * x) ** 2
    n, u, e, a = _moments(_maxwellians(g, n0, u0, T0), g)
    np.testing.assert_allclose(n, n0, rtol=1e-12)
    np.testing.assert_allclose(u, n0 * u0, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(a, 1.0 / T0, rtol=1e-11)


@pytest.mark.parametrize("kind", COLLS)
def test_collision_equals_assembled_stencil(kind):
    g = _grid(8, 24)
    rng = np.random.default_rng(3)
    f = rng.standard_normal((g.nx, g.nv))
    n = 1.0 + 0.1 * rng.random(g.nx)
    U = 0.2 * rng.standard_normal(g.nx)
    a = 1.0 + 0.5 * rng.random(g.nx)
    nu, diff, drag = 0.37, 1.3, 0.6
    y = O.collision_add(kind, f, np.zeros_like(f), n, U, a, 0.5, g, nu, diff, drag)
    lo, mid, hi = O.collision_stencil(kind, n, U, a, 0.5, g, nu, diff, drag)
    ref = np.zeros_like(f)
    ref[:, 1:-1] = lo[:, 1:-1] * f[:, :-2] + mid[:, 1:-1] * f[:, 1:-1] + hi[:, 1:-1] * f[:, 2:]
    assert np.abs(y[:, :2]).max() == 0.0 and np.abs(y[:, -2:]).max() == 0.0
    np.testing.assert_allclose(y, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("kind", COLLS)
def test_collision_conserves_mass_momentum_energy(kind):
    g = _grid(12, 48)
    rng = np.random.default_rng(4)
    f = _compact_random(g, rng)
    n, U, e, a = _moments(f, g)
    y = O.collision_add(kind, f, np.zeros_like(f), n, U, a, 1.0, g, 0.9, 1.0, 1.0)
    scale = np.abs(y).max() * g.nv * (g.vmax ** 2)
    assert np.abs(y.sum(axis=1)).max() <= 1e-13 * scale
    assert np.abs((y * g.v[None, :]).sum(axis=1)).max() <= 1e-13 * scale
    assert np.abs((y * g.v2[None, :]).sum(axis=1)).max() <= 1e-13 * scale
    # a wrong collision factor (the density in place of n^2) breaks energy conservation
    bad = O.collision_add(kind, f, np.zeros_like(f), n, U, a / n, 1.0, g, 0.9, 1.0, 1.0)
    assert np.abs((bad * g.v2[None, :]).sum(axis=1)).max() > 1e-6 * scale


@pytest.mark.parametrize("kind", COLLS)
def test_drifting_maxwellian_is_second_order_null(kind):
    errs = []
    for nv in (201, 401, 801):
        g = _grid(8, nv, vmax=10.0)
        x = np.arange(g.nx)
        f = _maxwellians(g, 1.0 + 0.1 * np.cos(x), 0.5 * np.sin(x), 1.0 + 0.2 * np.cos(x))
        n, U, e, a = _moments(f, g)
        y = O.collision_add(kind, f, np.zeros_like(f), n, U, a, 1.0, g, 1.0, 1.0, 1.0)
        # the drag and diffusion parts each are O(1); their sum is O(hv^2)
        part = O.collision_add(kind, f, np.zeros_like(f), n, U, a, 1.0, g, 1.0, 1.0, 0.0)
        assert np.abs(part).max() > 0.1
        errs.append(np.abs(y).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.5 < r1 < 4.5 and 3.5 < r2 < 4.5, errs


@pytest.mark.parametrize("kind", COLLS)
def test_diffusion_sign_is_dissipative(kind):
    """Y = F(f) + C(f): C's diffusion part enters the residual with the sign
    that makes the implicit step diffuse, i.e. sum f C_diff(f) > 0."""
    g = _grid(8, 48)
    rng = np.random.default_rng(5)
    f = _compact_random(g, rng)
    n, U, e, a = _moments(f, g)
    y = O.collision_add(kind, f, np.zeros_like(f), n, U, a, 1.0, g, 0.5, 1.0, 0.0)
    assert (f * y).sum() > 0


def _db_fields(g, rng):
    F = rng.standard_normal((g.nx, g.nv))
    H = rng.standard_normal((g.nx, g.nv))
    F[:, :6] = F[:, -6:] = 0.0
    H[:, :6] = H[:, -6:] = 0.0
    return F, H


@pytest.mark.parametrize("bkind", BRACKETS)
def test_double_bracket_identities(bkind):
    g = _grid(16, 48)
    rng = np.random.default_rng(6)
    F, H = _db_fields(g, rng)
    nu = 0.3
    B = O.bracket(bkind, F, H, g)
    b2 = (B * B).sum()
    ffh = O.double_bracket_function('ffh', bkind, F, H, np.zeros_like(F), 1.0, g, nu)
    fhh = O.double_bracket_function('fhh', bkind, F, H, np.zeros_like(F), 1.0, g, nu)
    tol = 1e-11 * nu * b2
    assert abs((H * ffh).sum() - nu * b2) <= tol
    assert abs((F * ffh).sum()) <= tol
    assert abs((F * fhh).sum() - nu * b2) <= tol
    assert abs((H * fhh).sum()) <= tol


@pytest.mark.parametrize("bkind", BRACKETS)
@pytest.mark.parametrize("kind", ['ffh', 'fhh'])
def test_double_bracket_jacobian_is_derivative(kind, bkind):
    """d/df of D(Fave), Fave = .5 fh + .5 f, in direction u equals the
    Jacobian term with factor .5; D is quadratic (ffh) or linear (fhh) in f."""
    g = _grid(12, 32)
    rng = np.random.default_rng(7)
    f, fh, u = rng.standard_normal((3, g.nx, g.nv))
    H = rng.standard_normal((g.nx, g.nv))
    nu, eps = 0.25, 1e-3

    def D(ff):
        return O.double_bracket_function(kind, bkind, 0.5 * fh + 0.5 * ff, H, np.zeros_like(ff), 1.0, g, nu)

    fd = (D(f + eps * u) - D(f - eps * u)) / (2 * eps)
    jac = O.double_bracket_jacobian(kind, bkind, u, 0.5 * fh + 0.5 * f, H, np.zeros_like(u), 0.5, g, nu)
    np.testing.assert_allclose(jac, fd, rtol=0, atol=1e-8 * np.abs(jac).max())


@pytest.mark.parametrize("dkind", ['collt', 'colle', 'ffh', 'fhh'])
def test_residual_with_dissipation_has_jacobian(dkind):
    """F(f) of the full function_solver is quadratic in f for fixed moments;
    its central differences equal jacobian_solver's action."""
    g = _grid(8, 32, vmax=5.0)
    rng = np.random.default_rng(8)
    f = _compact_random(g, rng)
    fh = _compact_random(g, rng)
    u = rng.standard_normal((g.nx, g.nv))
    H = O.h_average(O.kinetic_hamiltonian(g), *(0.1 * rng.standard_normal((4, g.nx))))
    n, U, e, a = _moments(f, g)
    nh, Uh, eh, ah = _moments(fh, g)
    mom = dict(np=n, up=U, ap=a, nh=nh, uh=Uh, ah=ah)
    diss = O.Dissipation(dkind, 0.2, 1.1, 0.9)
    eps = 1e-3

    def F(ff):
        return O.residual_diss('arakawaj4', ff, fh, H, g, diss, mom)

    fd = (F(f + eps * u) - F(f - eps * u)) / (2 * eps)
    jac = O.jacobian_diss('arakawaj4', u, 0.5 * fh + 0.5 * f, H, g, diss, mom)
    np.testing.assert_allclose(jac, fd, rtol=0, atol=1e-8 * np.abs(jac).max())
    # and the dissipation changes the residual (the term is not dropped)
    assert np.abs(F(f) - O.residual('arakawaj4', f, fh, H, g)).max() > 1e-6


def test_oracle_run_with_collisions_dissipates_energy_keeps_mass():
    """A collisional Landau step: mass is conserved, the implicit step still
    converges, and the distribution relaxes (sum f^2 decreases)."""
    g = _grid(16, 96, L=4 * np.pi, vmax=10.0, ht=0.1)
    X, V = np.meshgrid(g.x, g.v, indexing='ij')
    f0 = (1 + 0.05 * np.cos(0.5 * X)) * np.exp(-0.5 * (V - 1.0) ** 2) * (1 + 0.5 * np.exp(-2 * (V + 2) ** 2))
    run = O.OracleRun(g, f0, ksp_rtol=1e-12, snes_atol=1e-11,
                      dissipation=O.Dissipation('collt', 0.5, 1.0, 1.0))
    inv0 = O.invariants(run.fc, run.h0, run.h1c, run.h2c, run.phi, g, run.charge)
    l2 = [np.sum(run.fc ** 2)]
    for _ in range(3):
        e = run.step()
        assert e['picard'][-1]['reason'] > 0
        l2.append(np.sum(run.fc ** 2))
    inv = O.invariants(run.fc, run.h0, run.h1c, run.h2c, run.phi, g, run.charge)
    assert abs(inv['mass'] - inv0['mass']) <= 1e-9 * abs(inv0['mass'])
    assert l2[-1] < l2[0]
