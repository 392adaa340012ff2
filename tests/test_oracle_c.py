"""The C restatement (CPU baseline) against the NumPy oracle (CPU only)."""

import numpy as np
import pytest

from oracle import vlasov_oracle as O


@pytest.fixture(scope="module")
def cb():
    from oracle import build, cbaseline
    import os
    if not os.path.exists(cbaseline.LIB) or os.path.getmtime(cbaseline.LIB) < os.path.getmtime(build.SRC):
        build.build()
    return cbaseline


@pytest.mark.parametrize("nx,nv", [(16, 40), (9, 20), (50, 21)])
@pytest.mark.parametrize("kind", ['arakawaj1', 'arakawaj2', 'arakawaj4'])
def test_c_operators_match_oracle(cb, kind, nx, nv):
    g = O.Grid(nx, nv, 4 * np.pi, 8.0, ht=0.05)
    rng = np.random.default_rng(3)
    f0 = O.normalise(O.initial_distribution(g, lambda x, v: (1 + 0.2 * np.cos(0.5 * x)) * O.maxwellian(v)), g)
    run = cb.CRun(g, f0.T, bracket=kind)
    h1p, h1h, h2p, h2h = 0.1 * rng.standard_normal((4, g.nx))
    run.set_h(h1p, h1h, h2p, h2h)
    H = O.h_average(O.kinetic_hamiltonian(g), h1p, h1h, h2p, h2h)
    f, fh, u = rng.random((3, g.nx, g.nv))
    assert np.array_equal(run.residual(f.T, fh.T).T, O.residual(kind, f, fh, H, g))
    assert np.array_equal(run.jacobian(u.T).T, O.jacobian(kind, u, H, g))
    ref = O.Preconditioner(g).apply(u)
    got = run.precond(u.T).T
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("nx,nv", [(32, 64), (25, 48)])
def test_c_timesteps_match_oracle(cb, nx, nv):
    g = O.Grid(nx, nv, 4 * np.pi, 8.0, ht=0.05)
    prof = lambda x, v: (1 + 0.5 * np.cos(0.5 * x)) * O.maxwellian(v)
    orun = O.OracleRun(g, O.initial_distribution(g, prof))
    crun = cb.CRun(g, O.normalise(O.initial_distribution(g, prof), g).T)
    for _ in range(3):
        eo = orun.step()
        ec = crun.step()
        assert ec['picard'] == len(eo['picard'])
        assert abs(ec['gmres'] - sum(p['gmres_its'] for p in eo['picard'])) <= 2
        assert np.abs(crun.f().T - orun.fc).max() <= 1e-10 * np.abs(orun.fc).max()
        # the Picard loop converges to the same f even from poor linear
        # solves; the residual after the first solve (far above the rounding
        # floor, linear tolerance 1e-10) shows the solve itself
        r0 = eo['picard'][0]['residual']
        assert abs(ec['residuals'][0] - r0) <= 1e-8 * r0


@pytest.mark.parametrize("nx", [64, 128, 256, 512, 1024])
def test_c_precond_at_workload_lengths(cb, nx):
    # the x-lengths of the workloads (64 .. 1024; 1024 is the benchmark's),
    # each through the C restatement's own FFT, against numpy's rfft / irfft
    # in the oracle's preconditioner; few v-rows keep it fast
    g = O.Grid(nx, 16, 10 * np.pi, 10.0, ht=0.05)
    rng = np.random.default_rng(nx)
    f0 = O.normalise(O.initial_distribution(g, lambda x, v: (1 + 0.1 * np.cos(0.2 * x)) * O.maxwellian(v)), g)
    run = cb.CRun(g, f0.T)
    u = rng.standard_normal((g.nx, g.nv))
    ref = O.Preconditioner(g).apply(u)
    got = run.precond(u.T).T
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def _c_against_numpy(cb, g, prof, kind, steps):
    # every step: the same Picard count (or a tie at the stop decision, as
    # the GPU long runs allow) and f within 1e-12 of the NumPy oracle's
    from picard_ties import parting_kind
    orun = O.OracleRun(g, O.initial_distribution(g, prof), bracket=kind)
    crun = cb.CRun(g, O.normalise(O.initial_distribution(g, prof), g).T, bracket=kind)
    for _ in range(steps):
        eo, ec = orun.step(), crun.step()
        if parting_kind(ec['residuals'], [p['residual'] for p in eo['picard']], 1e-10) is not None:
            crun.set_f(orun.fc.T)
        assert np.abs(crun.f().T - orun.fc).max() <= 1e-12 * np.abs(orun.fc).max()


def test_c_linear_landau_full_run_matches_oracle(cb):
    # linear Landau 64x128, L = 4 pi, v in [-10, 10], amplitude 0.01, J4,
    # dt = 0.1: all 100 steps of the CPU-runnable workload
    g = O.Grid(64, 128, 4 * np.pi, 10.0, ht=0.1)
    _c_against_numpy(cb, g, lambda x, v: (1 + 0.01 * np.cos(0.5 * x)) * O.maxwellian(v), 'arakawaj4', 100)


@pytest.mark.parametrize("kind", ['arakawaj1', 'arakawaj2', 'arakawaj4'])
def test_c_nonlinear_landau_matches_oracle(cb, kind):
    # nonlinear Landau 256x512, v in [-8, 8], amplitude 0.5, dt = 0.05: the
    # first two steps at full size
    g = O.Grid(256, 512, 4 * np.pi, 8.0, ht=0.05)
    _c_against_numpy(cb, g, lambda x, v: (1 + 0.5 * np.cos(0.5 * x)) * O.maxwellian(v), kind, 2)
