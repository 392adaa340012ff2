"""The oracle's pins catch plausible mistakes (CPU only).

Each case plants one mistake in the oracle (a dropped term, a wrong sign, a
wrong stencil index, transposed operands, a wrong weight or normalisation)
in the Arakawa brackets, the residual and Jacobian, the kinetic
preconditioner's eigenvalues and bands, the Poisson stencil, solve and
right-hand side, the collision operators and factor, the double brackets,
the explicit steppers of the initial guess, GMRES, the Picard driver's
stop / revert rule, or the tangent-linear model of the adjoint (the last
seven edited into their source text); runs the independent pins of test_oracle_brackets.py and
test_oracle_solvers.py against the planted oracle; and requires that at
least two of them fail (every bracket case fails three or more; the
potential preconditioner's bands have one independent pin, the assembled
stencil tables, which each of their cases must fail).  Without the mistake
every pin passes."""

import inspect
import textwrap

import numpy as np
import pytest

import test_oracle_brackets as P
import test_oracle_dissipation as D
import test_oracle_guess as G
import test_oracle_tangent as T
from oracle import tangent_linear as TL
import test_oracle_solvers as SV
from oracle import vlasov_oracle as O

S = O._s
_J1, _J2 = O._j1_terms, O._j2_terms      # the unplanted terms


def _j1_without_jcp(x, h):
    jpp = (S(x, 1, 0) - S(x, -1, 0)) * (S(h, 0, 1) - S(h, 0, -1)) \
        - (S(x, 0, 1) - S(x, 0, -1)) * (S(h, 1, 0) - S(h, -1, 0))
    jpc = S(x, 1, 0) * (S(h, 1, 1) - S(h, 1, -1)) \
        - S(x, -1, 0) * (S(h, -1, 1) - S(h, -1, -1)) \
        - S(x, 0, 1) * (S(h, 1, 1) - S(h, -1, 1)) \
        + S(x, 0, -1) * (S(h, 1, -1) - S(h, -1, -1))
    return 1.5 * (jpp + jpc)          # dropped term, rescaled so sums still look right


def _j1_sign(x, h):
    # the last jpc term with the wrong sign
    return _J1(x, h) - 2 * S(x, 0, -1) * (S(h, 1, -1) - S(h, -1, -1))


def _j1_index(x, h):
    # S(h, 1, 1) read as S(h, 1, 0) in jpc's first term
    return _J1(x, h) - S(x, 1, 0) * (S(h, 1, 1) - S(h, 1, 0))


def _j2_index(x, h):
    # jcp's first term with (0, 2) read as (0, 1)
    return _J2(x, h) - S(x, 1, 1) * (S(h, 0, 2) - S(h, 0, 1))


def _j1_transposed(x, h):
    return _J1(h, x)


def _j4_weight(kind, x, h, y, factor, grid):
    af1 = grid.hx_inv * grid.hv_inv / 12.
    af2 = grid.hx_inv * grid.hv_inv / 24.
    inner = y[:, O.STENCIL:grid.nv - O.STENCIL]
    inner += factor * (O._j1_terms(x, h) * af1 - 0.5 * O._j2_terms(x, h) * af2)   # 2 rJ1 - rJ2 halved wrongly
    return y


def _j2_factor(kind, x, h, y, factor, grid):
    af = grid.hx_inv * grid.hv_inv / 12.                                        # 1/12 in place of 1/24
    y[:, O.STENCIL:grid.nv - O.STENCIL] += factor * O._j2_terms(x, h) * af
    return y


# (kinds the mistake reaches, what is replaced, the replacement)
CASES = {
    'j1 dropped jcp': (['arakawaj1', 'arakawaj4'], '_j1_terms', _j1_without_jcp),
    'j1 wrong sign': (['arakawaj1', 'arakawaj4'], '_j1_terms', _j1_sign),
    'j1 wrong index': (['arakawaj1', 'arakawaj4'], '_j1_terms', _j1_index),
    'j2 wrong index': (['arakawaj2', 'arakawaj4'], '_j2_terms', _j2_index),
    'j1 transposed operands': (['arakawaj1', 'arakawaj4'], '_j1_terms', _j1_transposed),
    'j4 wrong weight': (['arakawaj4'], 'bracket_add', _j4_weight),
    'j2 wrong normalisation': (['arakawaj2'], 'bracket_add', _j2_factor),
}


def _pins(kind):
    checks = [lambda: P.test_matrix_free_equals_stencil_table(kind), lambda: P.test_antisymmetry(kind),
              lambda: P.test_conservation(kind), lambda: P.test_closed_form_symbol(kind)]
    order = {'arakawaj1': 2, 'arakawaj2': 2, 'arakawaj4': 4}[kind]
    checks.append(lambda: P.test_consistency_order(kind, order))
    failed = 0
    for c in checks:
        try:
            c()
        except AssertionError:
            failed += 1
    return failed


@pytest.mark.parametrize("name", sorted(CASES))
def test_planted_mistake_fails_a_pin(name, monkeypatch):
    kinds, attr, repl = CASES[name]
    monkeypatch.setattr(O, attr, repl)
    for kind in kinds:
        assert _pins(kind) >= 2, (name, kind)


def test_unplanted_oracle_passes_every_pin():
    for kind in P.KINDS:
        assert _pins(kind) == 0, kind


# ------------------------------------------------ residual, P^-1 and Poisson
_RES, _JAC, _BANDS, _EIG = O.residual, O.jacobian, O.kinetic_bands, O.kinetic_eigenvalues
_PMM, _PEIG, _PRHS = O.poisson_matrix_mult, O.poisson_eigenvalues, O.poisson_rhs


def _res_fave(kind, f, fh, H, grid):
    y = grid.ht_inv * (f - fh)
    return O.bracket_add(kind, fh + 0.5 * f, H, y, 1.0, grid)          # Fave missing a 1/2


def _res_sign(kind, f, fh, H, grid):
    y = grid.ht_inv * (fh - f)                                          # Fder with the wrong sign
    return O.bracket_add(kind, 0.5 * fh + 0.5 * f, H, y, 1.0, grid)


def _jac_factor(kind, u, H, grid):
    return O.bracket_add(kind, u, H, grid.ht_inv * u, 1.0, grid)        # 1 in place of 1/2


def _bands_factor(grid, eigen_k):
    dm, d, dp = _BANDS(grid, eigen_k)                                   # a without its 1/2
    return 2 * dm, 2 * d - grid.ht_inv, 2 * dp


def _bands_swapped(grid, eigen_k):
    dm, d, dp = _BANDS(grid, eigen_k)                                   # sub / super-diagonal swapped
    return dp, d, dm


def _eig_conj(n):
    return np.conj(_EIG(n))                                             # exp(-2 pi i ...) for exp(+...)


def _pmm_coef(x, hx2_inv, laplace='cfd4'):
    y = _PMM(x, hx2_inv, laplace)
    return y - 2.0 * x * hx2_inv / 12. if laplace == 'cfd4' else y    # 54 read as 52


def _peig_coef(nx, hx, laplace='cfd4'):
    lam = _PEIG(nx, hx, laplace)
    th = 2.0 * np.pi * np.arange(nx) / nx
    if laplace == 'cfd4':
        lam = lam + 2. * np.cos(2. * th) / (12. * hx ** 2)              # 10 cos 2th read as 12 cos 2th
        lam[0] = 0.0
    return lam


def _prhs_sign(n, charge, nx):
    return -_PRHS(n, charge, nx)


_KINDS3 = ['arakawaj1', 'arakawaj2', 'arakawaj4']
_RES_PINS = [lambda k=k: P.test_residual_affine_and_jacobian(k) for k in _KINDS3] + \
            [lambda k=k: P.test_residual_structure(k) for k in _KINDS3]
_PC_PINS = [lambda: SV.test_eigenvalues_are_centered_difference_symbol(),
            lambda: SV.test_p_times_pinv_is_identity(8, 16),
            lambda: SV.test_preconditioner_is_half_j1_of_h0_plus_identity_over_dt(8),
            lambda: SV.test_preconditioner_is_half_j1_of_h0_plus_identity_over_dt(10),
            lambda: SV.test_odd_length_preconditioner_is_a_modewise_dense_solve()]
_POISSON_PINS = [lambda: SV.test_poisson_solve_vs_dense_lstsq('cfd4', 16),
                 lambda: SV.test_poisson_solve_vs_dense_lstsq('cfd4', 9),
                 lambda: SV.test_poisson_eigenvalues()]

SOLVER_CASES = {
    'residual Fave weight': ('residual', _res_fave, _RES_PINS),
    'residual time-derivative sign': ('residual', _res_sign, _RES_PINS),
    'jacobian bracket factor': ('jacobian', _jac_factor, _RES_PINS),
    'kinetic bands missing 1/2': ('kinetic_bands', _bands_factor, _PC_PINS),
    'kinetic bands swapped': ('kinetic_bands', _bands_swapped, _PC_PINS),
    'kinetic eigenvalues conjugated': ('kinetic_eigenvalues', _eig_conj, _PC_PINS),
    'poisson stencil coefficient': ('poisson_matrix_mult', _pmm_coef, _POISSON_PINS),
    'poisson eigenvalue coefficient': ('poisson_eigenvalues', _peig_coef, _POISSON_PINS),
    'poisson rhs sign': ('poisson_rhs', _prhs_sign, _POISSON_PINS),
}


def _failures(pins):
    failed = 0
    for c in pins:
        try:
            c()
        except AssertionError:
            failed += 1
    return failed


@pytest.mark.parametrize("name", sorted(SOLVER_CASES))
def test_planted_solver_mistake_fails_a_pin(name, monkeypatch):
    attr, repl, pins = SOLVER_CASES[name]
    monkeypatch.setattr(O, attr, repl)
    assert _failures(pins) >= 2, name


def test_unplanted_solver_oracle_passes_every_pin():
    assert _failures(_RES_PINS + _PC_PINS + _POISSON_PINS) == 0


# ------------------------------------------------------------------- GMRES
# (text in O.gmres, its replacement)
GMRES_CASES = {
    'previous rotation sign': ('H[jj + 1, it] = cs[jj] * H[jj + 1, it] - sn[jj] * t',
                               'H[jj + 1, it] = cs[jj] * H[jj + 1, it] + sn[jj] * t'),
    'right-hand-side rotation sign': ('g[it + 1] = -(sn[it] * g[it])', 'g[it + 1] = (sn[it] * g[it])'),
    'back substitution drops a term': ('for jj in range(k + 1, it):', 'for jj in range(k + 2, it):'),
    'CGS skips the newest vector': ('for jj in range(it + 1):', 'for jj in range(it):'),
    'restart residual sign': ('r = b.copy() if first else b - apply_A(x)', 'r = b.copy() if first else b + apply_A(x)'),
    'update from the wrong basis vector': ('upd = upd + y[jj] * V[jj]', 'upd = upd + y[jj] * V[jj + 1]'),
}
_GMRES_PINS = [lambda r=r: SV.test_gmres_vs_dense_solve(r) for r in (50, 10, 3)] + \
              [SV.test_gmres_full_restart_equivalence, SV.test_gmres_edge_cases]


@pytest.mark.parametrize("name", sorted(GMRES_CASES))
def test_planted_gmres_mistake_fails_a_pin(name, monkeypatch):
    old, new = GMRES_CASES[name]
    src = inspect.getsource(O.gmres)
    assert src.count(old) == 1, name
    ns = dict(vars(O))
    exec(src.replace(old, new), ns)
    monkeypatch.setattr(O, 'gmres', ns['gmres'])
    with np.errstate(all='ignore'):
        assert _failures(_GMRES_PINS) >= 2, name


def test_unplanted_gmres_passes_every_pin():
    assert _failures(_GMRES_PINS) == 0


# ------------------------------------------------------ collision operators
# (text in O.collision_add, its replacement, the operators it reaches)
COLLISION_CASES = {
    'drag without its 1/2': ('grid.hv_inv * 0.5', 'grid.hv_inv', ['collt', 'colle']),
    'diffusion over hv, not hv^2': ('coll_diff * grid.hv2_inv', 'coll_diff * grid.hv_inv', ['collt', 'colle']),
    'collT drag sign': ('drag = ((vp - u) * fp - (vm - u) * fm) * aa',
                        'drag = ((vp - u) * fp + (vm - u) * fm) * aa', ['collt']),
    'velocity density not divided by n': ('u = (U / n)[:, None]', 'u = U[:, None]', ['collt', 'colle']),
    'rows j+1 and j-1 swapped': ('fp, f0, fm = f[', 'fm, f0, fp = f[', ['collt', 'colle']),
    'collE diffusion times a': ('diff = ((fp - 2. * f0) + fm) / aa', 'diff = ((fp - 2. * f0) + fm) * aa', ['colle']),
}


def _collision_pins(kind):
    return [lambda: D.test_collision_equals_assembled_stencil(kind),
            lambda: D.test_collision_conserves_mass_momentum_energy(kind),
            lambda: D.test_drifting_maxwellian_is_second_order_null(kind),
            lambda: D.test_diffusion_sign_is_dissipative(kind)]


@pytest.mark.parametrize("name", sorted(COLLISION_CASES))
def test_planted_collision_mistake_fails_a_pin(name, monkeypatch):
    old, new, kinds = COLLISION_CASES[name]
    src = inspect.getsource(O.collision_add)
    assert src.count(old) == 1, name
    ns = dict(vars(O))
    exec(src.replace(old, new), ns)
    monkeypatch.setattr(O, 'collision_add', ns['collision_add'])
    for kind in kinds:
        with np.errstate(all='ignore'):
            assert _failures(_collision_pins(kind)) >= 2, (name, kind)


def test_planted_collision_factor_mistake_fails_a_pin(monkeypatch):
    # n / (n e - u u): the density in place of n^2
    monkeypatch.setattr(O, 'collision_factor', lambda n, u, e: n / (n * e - u * u))
    pins = [D.test_collision_factor_is_inverse_temperature] + _collision_pins('collt') + _collision_pins('colle')
    assert _failures(pins) >= 2


def test_unplanted_collisions_pass_every_pin():
    assert _failures([D.test_collision_factor_is_inverse_temperature] + _collision_pins('collt')
                     + _collision_pins('colle')) == 0


# ---------------------------------------- double brackets and explicit steppers
def _planted(fn, old, new):
    src = inspect.getsource(getattr(O, fn))
    assert src.count(old) == 1, (fn, old)
    ns = dict(vars(O))
    exec(src.replace(old, new), ns)
    return ns[fn]


DB = '-coll_freq * factor, grid)'
DOUBLE_BRACKET_CASES = {
    'FFH arguments swapped': ('double_bracket_function', 'return bracket_add(bkind, fave, b, y, ' + DB,
                              'return bracket_add(bkind, b, fave, y, ' + DB),
    'FHH sign': ('double_bracket_function', 'return bracket_add(bkind, b, H, y, ' + DB,
                 'return bracket_add(bkind, b, H, y, coll_freq * factor, grid)'),
    'FFH Jacobian drops a term': ('double_bracket_jacobian', 'return bracket_add(bkind, fave, b, y, ' + DB, 'return y'),
    'FHH Jacobian arguments swapped': ('double_bracket_jacobian', 'return bracket_add(bkind, b, H, y, ' + DB,
                                       'return bracket_add(bkind, H, b, y, ' + DB),
}
_DB_PINS = [lambda b=b: D.test_double_bracket_identities(b) for b in D.BRACKETS] + \
           [lambda b=b, k=k: D.test_double_bracket_jacobian_is_derivative(k, b)
            for b in D.BRACKETS for k in ('ffh', 'fhh')] + \
           [lambda k=k: D.test_residual_with_dissipation_has_jacobian(k) for k in ('ffh', 'fhh')]

STEPPER_CASES = {
    'RK4 (1/6) weights over 5': ('rk4_16', '(t1 + 2. * t2 + 2. * t3 + t4) / 6.', '(t1 + 2. * t2 + 2. * t3 + t4) / 5.'),
    'RK4 (1/6) third stage from t1': ('rk4_16', 't3 = tendency(kind, x + 0.5 * grid.ht / float(niter) * t2, h, grid)',
                                      't3 = tendency(kind, x + 0.5 * grid.ht / float(niter) * t1, h, grid)'),
    'RK4 (3/8) third stage sign': ('rk4_18', 't3 = tendency(kind, x - 1. / 3. * grid.ht',
                                   't3 = tendency(kind, x + 1. / 3. * grid.ht'),
    'RK4 (3/8) weight': ('rk4_18', '(t1 + 3. * t2 + 3. * t3 + t4) / 8.', '(t1 + 2. * t2 + 3. * t3 + t4) / 8.'),
    'Forest-Ruth c2 sign': ('symplectic4_coefficients', 'c2 = c1 * (1. - fac2)', 'c2 = c1 * (fac2 - 1.)'),
    'kinetic substep without 1/niter': ('symplectic_kinetic', 'factor * grid.ht / float(niter)', 'factor * grid.ht'),
    'potential substep sign': ('symplectic_potential', 'x + factor * grid.ht / float(niter) * tendency',
                               'x - factor * grid.ht / float(niter) * tendency'),
}
_STEPPER_PINS = [G.test_symplectic4_coefficients_are_forest_ruth] + \
                [lambda s=s: G.test_symplectic_is_product_of_euler_substeps(s) for s in ('symplectic2', 'symplectic4')] + \
                [lambda n=n: G.test_symplectic_substep_divides_the_step(n) for n in (1, 2, 3)] + \
                [lambda s=s, n=n: G.test_rk4_is_the_fourth_order_taylor_polynomial(s, n)
                 for s in ('rk4_16', 'rk4_18') for n in (1, 3)] + \
                [lambda s=s, r=r: G.test_one_step_error_order(s, r)
                 for s, r in (('rk4_16', 32.0), ('rk4_18', 32.0), ('symplectic2', 4.0), ('symplectic4', 4.0))] + \
                [lambda s=s: G.test_mass_is_conserved(s) for s in ('rk4_16', 'rk4_18', 'symplectic2', 'symplectic4')]


@pytest.mark.parametrize("name", sorted(DOUBLE_BRACKET_CASES))
def test_planted_double_bracket_mistake_fails_a_pin(name, monkeypatch):
    fn, old, new = DOUBLE_BRACKET_CASES[name]
    monkeypatch.setattr(O, fn, _planted(fn, old, new))
    assert _failures(_DB_PINS) >= 2, name


@pytest.mark.parametrize("name", sorted(STEPPER_CASES))
def test_planted_stepper_mistake_fails_a_pin(name, monkeypatch):
    fn, old, new = STEPPER_CASES[name]
    monkeypatch.setattr(O, fn, _planted(fn, old, new))
    assert _failures(_STEPPER_PINS) >= 2, name


def test_unplanted_double_brackets_and_steppers_pass_every_pin():
    assert _failures(_DB_PINS + _STEPPER_PINS) == 0


# ----------------------------------------------------------- Picard driver
DRIVER_CASES = {
    'a first-iteration rise stops the loop': ('(pred > prev and i > 1) or', '(pred > prev) or'),
    'snes_atol itself stops the loop': ('pred < self.snes_atol', 'pred <= self.snes_atol'),
    'iteration cap off by one': ('i >= self.snes_max_iter', 'i > self.snes_max_iter'),
    'no revert at the first iteration': ('            if pred > prev:\n', '            if pred > prev and i > 1:\n'),
    'the revert keeps the iterate': ('self.fc = fl\n', 'pass\n'),
}
_DRIVER_PINS = [lambda c=c, e=e: SV.test_driver_stop_and_revert_rule(c, e) for c, e in SV.PRESCRIBED] + \
               [SV.test_driver_max_iter_branch, SV.test_driver_revert_branch]


@pytest.mark.parametrize("name", sorted(DRIVER_CASES))
def test_planted_driver_mistake_fails_a_pin(name, monkeypatch):
    old, new = DRIVER_CASES[name]
    src = textwrap.dedent(inspect.getsource(O.OracleRun.step))
    assert src.count(old) == 1, name
    ns = dict(vars(O))
    exec(src.replace(old, new), ns)
    monkeypatch.setattr(O.OracleRun, 'step', ns['step'])
    assert _failures(_DRIVER_PINS) >= 2, name


def test_unplanted_driver_passes_every_pin():
    assert _failures(_DRIVER_PINS) == 0


# ------------------------------------------------ potential preconditioner
_PBANDS = O.potential_bands


def _pb(change):
    return lambda grid, phi, e: change(grid, *_PBANDS(grid, phi, e))


POTENTIAL_CASES = {
    'bands without the 1/2': _pb(lambda g, dm, d, dp: (2 * dm, 2 * d - g.ht_inv, 2 * dp)),
    'sub- and super-diagonal swapped': _pb(lambda g, dm, d, dp: (dp, d, dm)),
    'phi terms with the wrong sign': _pb(lambda g, dm, d, dp: (-dm, 2 * g.ht_inv - d, -dp)),
    'main-diagonal weight 1 for 2': _pb(lambda g, dm, d, dp: (dm, 0.5 * (d - g.ht_inv) + g.ht_inv, dp)),
}
_POTENTIAL_PINS = [SV.test_potential_preconditioner_is_minus_half_j1_of_phi_plus_identity_over_dt,
                   SV.test_potential_preconditioner_variants_and_inverse,
                   SV.test_driver_with_potential_preconditioner]


@pytest.mark.parametrize("name", sorted(POTENTIAL_CASES))
def test_planted_potential_band_mistake_fails_the_stencil_pin(name, monkeypatch):
    monkeypatch.setattr(O, 'potential_bands', POTENTIAL_CASES[name])
    with pytest.raises(AssertionError):
        SV.test_potential_preconditioner_is_minus_half_j1_of_phi_plus_identity_over_dt()


def test_unplanted_potential_preconditioner_passes_every_pin():
    assert _failures(_POTENTIAL_PINS) == 0


# ------------------------------------------------------- tangent-linear model
TANGENT_CASES = {
    'field coupling with 1 for 1/2': ('O.bracket_add(kind, fbar, hu, y, 0.5, grid)', 'O.bracket_add(kind, fbar, hu, y, 1.0, grid)'),
    'field coupling dropped': ('O.bracket_add(kind, fbar, hu, y, 0.5, grid)', 'pass'),
    'A0 with the sign of A': ("sign = 1.0 if which == 'A' else -1.0", 'sign = 1.0'),
    'midpoint taken at f1': ('fbar = 0.5 * f0 + 0.5 * f1', 'fbar = f1'),
}
_TANGENT_PINS = [lambda k=k: T.test_linearisation_is_the_central_difference(k) for k in T.KINDS] + \
                [T.test_tangent_linear_matches_picard_differences]


@pytest.mark.parametrize("name", sorted(TANGENT_CASES))
def test_planted_tangent_mistake_fails_a_pin(name, monkeypatch):
    old, new = TANGENT_CASES[name]
    src = inspect.getsource(TL.linear_apply)
    assert src.count(old) == 1, name
    ns = dict(vars(TL))
    exec(src.replace(old, new), ns)
    monkeypatch.setattr(TL, 'linear_apply', ns['linear_apply'])
    assert _failures(_TANGENT_PINS) >= 2, name


def test_unplanted_tangent_passes_every_pin():
    assert _failures(_TANGENT_PINS) == 0
