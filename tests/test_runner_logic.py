"""The product runner's Picard stop / revert / abort rules
(run_nonlinear_matrixfree_split_mp.py:140-168), forced from the host on the
CPU: Runscript.step() runs as shipped, with its device calls (history,
moments, the SNES solve) replaced by host stand-ins and the residual
sequence prescribed.  Covers the branches the GPU runs rarely or never
reach: the abort (residual > 10 -> ResidualTooLarge, where the reference
calls sys.exit(1)), the revert of an increased residual, the iteration cap,
and convergence below snes_atol."""

import pytest
import torch

from vivlasov1d_amd.runner import ResidualTooLarge, Runscript


class _KSP:
    precision = 'double'

    def getIterationNumber(self):
        return 7

    def getConvergedReason(self):
        return 2


class _SNES:
    def __init__(self, run):
        self.run = run
        self.ksp = _KSP()

    def solve(self, b, x):
        x.add_(1.0)          # the Newton update moves f; a revert must undo it
        self.run.solves += 1

    def getKSP(self):
        return self.ksp


class _Solver:
    def update_previous(self, f):
        pass


class _Grid:
    ht = 0.1


def _run(residuals, snes_atol=1e-10, snes_max_iter=20):
    run = object.__new__(Runscript)
    run.itime = 0
    run.grid = _Grid()
    run.external = None
    run.initial_guess_method = None
    run.vlasov_solver = _Solver()
    run.fc = torch.zeros(4, dtype=torch.float64)
    run.fl = torch.zeros(4, dtype=torch.float64)
    run.snes = _SNES(run)
    run.snes_atol, run.snes_max_iter = snes_atol, snes_max_iter
    run.history, run.log, run.output, run.verbose, run.nsave = [], None, None, False, 1
    run.solves = 0
    run.moments_calls = 0
    seq = iter(residuals)
    run.make_history = lambda: None
    run.calculate_residual = lambda: next(seq)

    def moments():
        run.moments_calls += 1
    run.calculate_moments = moments
    return run


def test_converges_below_snes_atol():
    run = _run([1.0, 1e-3, 1e-7, 5e-11])
    rec = run.step()
    assert len(rec['picard']) == 3 and rec['final'] == 5e-11 and 'reverted' not in rec
    assert float(run.fc[0]) == 3.0


def test_increase_reverts_the_last_iterate():
    run = _run([1.0, 1e-3, 2e-3])
    rec = run.step()
    assert len(rec['picard']) == 2 and rec['reverted']
    assert float(run.fc[0]) == 1.0          # the second update is undone
    assert run.moments_calls == 3           # after each solve, and after the revert


def test_first_iteration_increase_does_not_stop():
    run = _run([1.0, 2.0, 1e-11])
    rec = run.step()
    assert len(rec['picard']) == 2 and 'reverted' not in rec


def test_first_iteration_increase_reverts_when_stopping_below_atol():
    # the stop test's "i > 1" guards only the rise; a first iterate that
    # rose and lies below snes_atol still stops the loop, and is undone
    run = _run([2e-11, 5e-11, 1e-12])
    rec = run.step()
    assert len(rec['picard']) == 1 and rec['reverted'] and rec['final'] == 5e-11
    assert float(run.fc[0]) == 0.0 and run.moments_calls == 2


def test_first_iteration_increase_reverts_at_the_cap():
    run = _run([1.0, 2.0, 1e-3], snes_max_iter=1)
    rec = run.step()
    assert len(rec['picard']) == 1 and rec['reverted'] and run.solves == 1
    assert float(run.fc[0]) == 0.0


def test_iteration_cap():
    run = _run([1.0] + [0.5 ** k for k in range(1, 30)], snes_max_iter=5)
    rec = run.step()
    assert len(rec['picard']) == 5 and run.solves == 5


def test_residual_too_large_aborts():
    run = _run([50.0, 20.0, 30.0])
    with pytest.raises(ResidualTooLarge):
        run.step()
    # the step is recorded before the abort, with the revert applied
    assert run.history[-1]['final'] == 30.0 and run.history[-1]['reverted']
    assert float(run.fc[0]) == 1.0
