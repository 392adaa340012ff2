"""GPU parity of the explicit initial guess against the oracle.

Bars: the symplectic substeps and the RK4 steps are bit-identical to the
oracle (bracket store mode with factor -1, stage combinations in NumPy's
evaluation order); Picard runs with each guess reproduce the oracle run's
guess decisions and Picard counts, and f agrees to 1e-9 relative.
"""

import math

import numpy as np
import pytest

from oracle import vlasov_oracle as O

pytestmark = pytest.mark.gpu


def _setup(nx, nv, kind, cuda, seed=0, ht=0.05):
    import torch
    from vivlasov1d_amd.core import Grid as PGrid
    from vivlasov1d_amd.solvers import Context
    g = O.Grid(nx, nv, 4 * np.pi, 6.0, ht=ht)
    pg = PGrid(nx, nv, 4 * np.pi, 6.0, ht=ht)
    ctx = Context(pg, bracket=kind, preconditioner='tensor_kinetic')
    rng = np.random.default_rng(seed)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    f = rng.random((nx, nv))
    h0 = O.kinetic_hamiltonian(g)
    h1, h2 = 0.2 * rng.standard_normal(nx), 0.05 * rng.standard_normal(nx)
    return g, ctx, T, f, h0, h1, h2


@pytest.mark.parametrize("kind", ['arakawaj1', 'arakawaj2', 'arakawaj4'])
@pytest.mark.parametrize("niter", [1, 3])
def test_steppers_bitwise(kind, niter, cuda):
    print("case", kind, niter, flush=True)
    g, ctx, T, f, h0, h1, h2 = _setup(64, 96, kind, cuda, seed=niter)
    H0 = np.broadcast_to(h0[None, :], (g.nx, g.nv))
    H12 = np.broadcast_to((h1 + h2)[:, None], (g.nx, g.nv))
    H = (H0 + h1[:, None]) + h2[:, None]
    h0d, h1d, h2d = T(h0), T(h1), T(h2)
    c, d = O.symplectic4_coefficients()
    for factor in (0.5, c[1]):
        x = T(f.T)
        ctx.symplectic_kinetic(x, h0d, factor, niter)
        assert np.array_equal(x.cpu().numpy().T, O.symplectic_kinetic(kind, f, H0, factor, g, niter))
    for factor in (1.0, d[1]):
        x = T(f.T)
        ctx.symplectic_potential(x, h1d, h2d, factor, niter)
        assert np.array_equal(x.cpu().numpy().T, O.symplectic_potential(kind, f, H12, factor, g, niter))
    for rule, ref in (('rk4_16', O.rk4_16), ('rk4_18', O.rk4_18)):
        x = T(f.T)
        ctx.rk4(x, h0d, h1d, h2d, rule, niter)
        got, want = x.cpu().numpy().T, ref(kind, f, H, g, niter)
        assert np.array_equal(got, want), (rule, np.abs(got - want).max())


def _cfg(guess, niter, nx, nv, L, vmax, ht):
    from vivlasov1d_amd.core import Config
    return Config({'grid': {'nx': nx, 'nv': nv, 'L': L, 'vmax': vmax, 'ht': ht, 'nt': 3},
                   'solver': {'poisson_bracket': 'ArakawaJ4', 'laplace_operator': 'CFD4',
                              'preconditioner_type': 'tensor', 'preconditioner_scheme': 'tensor_kinetic',
                              'petsc_ksp_rtol': 1e-10, 'petsc_snes_atol': 1e-10, 'petsc_snes_max_iter': 20,
                              'initial_guess': guess, 'initial_iter': niter},
                   'initial_data': {'charge': -1.0}})


@pytest.mark.parametrize("guess,niter", [('symplectic2', 1), ('symplectic4', 2), ('rk4', 1), ('rk4_16', 1)])
def test_picard_with_initial_guess_matches_oracle(guess, niter, cuda):
    from vivlasov1d_amd.runner import Runscript
    print("case", guess, niter, flush=True)
    nx, nv, L, vmax, ht = 32, 64, 4 * np.pi, 8.0, 0.1
    prof = lambda x, v: (1 + 0.3 * np.cos(0.5 * x)) * np.exp(-0.5 * (v - 0.5) ** 2) / math.sqrt(2 * np.pi)
    g = O.Grid(nx, nv, L, vmax, ht=ht)
    orun = O.OracleRun(g, O.initial_distribution(g, prof), ksp_rtol=1e-10, snes_atol=1e-10,
                       initial_guess=guess, initial_iter=niter)
    run = Runscript(_cfg(guess, niter, nx, nv, L, vmax, ht), profile=prof)
    for s in range(3):
        eo = orun.step()
        eg = run.step()
        (po, io), (pg, ig) = eo['guess'], eg['guess']
        print(s, po, pg, io, ig, len(eo['picard']), len(eg['picard']), flush=True)
        assert abs(pg - po) <= 1e-9 * po + 1e-13 and abs(ig - io) <= 1e-9 * io
        assert (ig > pg) == (io > po)
        assert len(eg['picard']) == len(eo['picard'])
        f = run.fc.cpu().numpy().T
        assert np.abs(f - orun.fc).max() <= 1e-9 * np.abs(orun.fc).max()


def test_unknown_initial_guess_is_an_error(cuda):
    from vivlasov1d_amd import _lib
    from vivlasov1d_amd.runner import Runscript
    prof = lambda x, v: np.exp(-0.5 * v ** 2)
    with pytest.raises(_lib.VVError, match="initial guess"):
        Runscript(_cfg('gear4', 1, 16, 32, 2 * np.pi, 6.0, 0.1), profile=prof)
    with pytest.raises(_lib.VVError, match="Gear history"):
        Runscript(_cfg('gear', 1, 16, 32, 2 * np.pi, 6.0, 0.1), profile=prof)
