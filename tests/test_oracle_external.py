"""Pins of the oracle's external potential (run_base.py:514-525, 707-747;
run_nonlinear_matrixfree_split_mp.py:124-125), CPU only.

* h2 after step s is the closed form external(x, s dt) minus its mean;
* h2 enters the step: a driven run departs from the undriven one by far
  more than the solver tolerance, and the departure flips sign with the
  drive (to first order in a weak drive);
* the driven run still conserves mass (every bracket term is a divergence).
"""

import math

import numpy as np

from oracle import vlasov_oracle as O


def _landau(g):
    return O.initial_distribution(g, lambda x, v: (1 + 0.1 * np.cos(0.5 * x)) * O.maxwellian(v))


def _drive(amp):
    return lambda x, t: amp * math.cos(0.5 * x - 0.3 * t)


def test_h2_is_the_mean_free_external_potential():
    g = O.Grid(16, 32, 4 * np.pi, 6.0, ht=0.1)
    run = O.OracleRun(g, _landau(g), external=_drive(0.05), preconditioner=None, snes_atol=1e-11)
    for s in range(3):
        t = s * g.ht
        p = 0.05 * np.cos(0.5 * g.x - 0.3 * t)
        p = p - p.mean()
        assert np.abs(run.h2c - p).max() <= 1e-16 * 16
        assert np.abs(run.pc_ext - p).max() <= 1e-16 * 16
        run.step()
    assert run.itime == 3
    # the history potential is the previous step's
    p2 = 0.05 * np.cos(0.5 * g.x - 0.3 * 2 * g.ht)
    assert np.abs(run.h2h - (p2 - p2.mean())).max() <= 1e-16 * 16


def test_drive_changes_the_step_and_conserves_mass():
    g = O.Grid(16, 48, 4 * np.pi, 10.0, ht=0.1)     # f ~ e^-50 at the masked v-boundary rows
    runs = {a: O.OracleRun(g, _landau(g), external=(_drive(a) if a else None), snes_atol=1e-12,
                           ksp_rtol=1e-12) for a in (0.0, 1e-3, -1e-3)}
    m0 = O.density(runs[0.0].fc, g).sum()
    for _ in range(3):
        for r in runs.values():
            r.step()
    d_plus = runs[1e-3].fc - runs[0.0].fc
    d_minus = runs[-1e-3].fc - runs[0.0].fc
    scale = np.abs(runs[0.0].fc).max()
    assert np.abs(d_plus).max() >= 1e-6 * scale
    # linear response: the departure is odd in the drive amplitude
    assert np.abs(d_plus + d_minus).max() <= 1e-2 * np.abs(d_plus).max()
    for r in runs.values():
        assert abs(O.density(r.fc, g).sum() - m0) <= 1e-12 * m0


def test_identity_preconditioner_run_converges():
    g = O.Grid(16, 32, 4 * np.pi, 6.0, ht=0.1)
    run = O.OracleRun(g, _landau(g), preconditioner=None, snes_atol=1e-11)
    e = run.step()
    assert e['picard'][-1]['residual'] < 1e-11
    assert e['picard'][0]['gmres_its'] > 3
