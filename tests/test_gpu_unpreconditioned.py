"""Runs without a preconditioner (preconditioner_type None, the default of
the reference's examples landau_linear.cfg, bump_on_tail.cfg, keen.cfg ...):
GMRES on J itself, SNES ksponly with F unpreconditioned.

Bars: GMRES converged reason equal, iteration count within 1 and solution
within 1e-8 relative of the oracle's GMRES; mixed GMRES: true FP64 residual
within the tolerance and solution within 1e-8; Picard steps: the same Picard
count and f within 1e-9 relative of the oracle run.
"""

import math

import numpy as np
import pytest

from oracle import vlasov_oracle as O

pytestmark = pytest.mark.gpu


def _nat(a):
    return np.ascontiguousarray(a.T)


@pytest.mark.parametrize("nx,nv", [(32, 64), (51, 101)])
def test_gmres_on_j_matches_oracle(nx, nv, cuda):
    import torch
    from vivlasov1d_amd.core import Grid as PGrid
    from vivlasov1d_amd.solvers import Context
    print("case", nx, nv, flush=True)
    rng = np.random.default_rng(nx)
    g = O.Grid(nx, nv, 4 * np.pi, 8.0, ht=0.1)
    ctx = Context(PGrid(nx, nv, 4 * np.pi, 8.0, ht=0.1), bracket='arakawaj4', preconditioner=None)
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    h0 = O.kinetic_hamiltonian(g)
    h1p, h1h = 0.1 * rng.standard_normal((2, nx))
    z = np.zeros(nx)
    ctx.set_hamiltonian(T(h0), T(h1p), T(h1h), T(z), T(z))
    H = O.h_average(h0, h1p, h1h, z, z)
    f = rng.random((nx, nv))
    f[:, :2] = 0
    f[:, -2:] = 0
    fh = f + 1e-3 * rng.standard_normal((nx, nv))
    b = O.residual('arakawaj4', f, fh, H, g)
    A = lambda v: O.jacobian('arakawaj4', v, H, g)
    x_ref, its_ref, reason_ref = O.gmres(A, b, rtol=1e-10, atol=1e-50, max_it=1000, restart=30)
    xd = ctx.new_field()
    res = ctx.ksp_solve(T(_nat(b)), xd, 1e-10, 1e-50, 1e5, 1000)
    assert res.reason == reason_ref, (res.reason, reason_ref)
    assert abs(res.its - its_ref) <= 1, (res.its, its_ref)
    x = xd.cpu().numpy().T
    assert np.abs(x - x_ref).max() <= 1e-8 * np.abs(x_ref).max()
    # mixed precision (FP32 cycles on J, FP64 refinement)
    xm = ctx.new_field()
    resm, cycles = ctx.ksp_solve_mixed(T(_nat(b)), xm, 1e-10, 1e-50, 1e5, 1000, inner_rtol=1e-5)
    assert resm.reason in (O.KSP_CONVERGED_RTOL, O.KSP_CONVERGED_ATOL), resm.reason
    xm = xm.cpu().numpy().T
    assert np.linalg.norm(b - A(xm)) <= 1e-10 * np.linalg.norm(b) * (1 + 1e-6)
    assert np.abs(xm - x_ref).max() <= 1e-8 * np.abs(x_ref).max()


LANDAU_LINEAR = "import numpy as np\ndef density(x, L):\n    return 1. + 0.01 * np.cos(0.5 * (x - 0.5 * L))\n"


def test_landau_linear_config_matches_oracle(cuda, tmp_path, monkeypatch):
    """examples/landau_linear.cfg as written: 100 x 201, v in [-10, 10],
    dt 0.1, ksp_rtol 1e-20 / ksp_atol 1e-11 / 100 iterations, snes_atol
    1e-10, no preconditioner, density_python landau_linear (restated), 3
    steps."""
    from vivlasov1d_amd.core import Config, initial_data
    from vivlasov1d_amd.runner import Runscript
    (tmp_path / "vv_landau_linear.py").write_text(LANDAU_LINEAR)
    monkeypatch.syspath_prepend(str(tmp_path))
    cfg = Config({'grid': {'nt': 100, 'nx': 100, 'nv': 201, 'ht': 0.1, 'L': 12.566370614359172,
                           'vmin': -10.0, 'vmax': 10.0},
                  'solver': {'petsc_snes_rtol': 1e-20, 'petsc_snes_atol': 1e-10, 'petsc_snes_stol': 1e-14,
                             'petsc_snes_max_iter': 20, 'petsc_ksp_rtol': 1e-20, 'petsc_ksp_atol': 1e-11,
                             'petsc_ksp_max_iter': 100, 'coll_freq': 0.},
                  'initial_data': {'charge': -1.0, 'temperature': 1.0, 'density_python': 'vv_landau_linear'}})
    run = Runscript(cfg)
    g = O.Grid(100, 201, 12.566370614359172, 10.0, -10.0, ht=0.1)
    orun = O.OracleRun(g, initial_data(cfg, run.grid).T, charge=-1.0, preconditioner=None, ksp_rtol=1e-20,
                       ksp_atol=1e-11, ksp_max_it=100, snes_atol=1e-10)
    assert np.abs(run.fc.cpu().numpy().T - orun.fc).max() <= 1e-14 * np.abs(orun.fc).max()
    for s in range(3):
        eo = orun.step()
        eg = run.step()
        print(s, [p['gmres_its'] for p in eo['picard']], [p['gmres'] for p in eg['picard']], flush=True)
        assert len(eg['picard']) == len(eo['picard'])
        f = run.fc.cpu().numpy().T
        assert np.abs(f - orun.fc).max() <= 1e-9 * np.abs(orun.fc).max(), s
