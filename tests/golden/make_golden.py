"""Write the golden fixtures of tests/golden/ from the NumPy oracle.

The reference ships no test vectors (no tests, fixtures or stored outputs)
and cannot be built here (petsc4py, PETSc, FFTW absent), so the fixtures are
produced by the oracle after it has passed its pins (tests/test_oracle_*.py).
They anchor regressions of the oracle and give the GPU tests data that does
not depend on the host's numpy/scipy.

Run:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import vlasov_oracle as O  # noqa: E402

KINDS = ['arakawaj1', 'arakawaj2', 'arakawaj4']


def brackets():
    g = O.Grid(16, 32, 4 * np.pi, 8.0, ht=0.05)
    rng = np.random.default_rng(2024)
    f = rng.random((g.nx, g.nv))
    fh = f + 0.01 * rng.standard_normal((g.nx, g.nv))
    u = rng.standard_normal((g.nx, g.nv))
    h1p, h1h, h2p, h2h = 0.1 * rng.standard_normal((4, g.nx))
    H = O.h_average(O.kinetic_hamiltonian(g), h1p, h1h, h2p, h2h)
    out = dict(nx=g.nx, nv=g.nv, L=g.L, vmax=g.vmax, ht=g.ht, f=f, fh=fh, u=u, h1p=h1p, h1h=h1h, h2p=h2p, h2h=h2h)
    for k in KINDS:
        out['res_' + k] = O.residual(k, f, fh, H, g)
        out['jac_' + k] = O.jacobian(k, u, H, g)
    out['precond_x'] = u
    out['precond_y'] = O.Preconditioner(g, 'lapack').apply(u)
    np.savez_compressed(os.path.join(HERE, "operators_16x32.npz"), **out)


def poisson():
    nx, L = 32, 10 * np.pi
    hx = L / nx
    rng = np.random.default_rng(7)
    n = 1 + 0.1 * rng.standard_normal(nx)
    out = dict(nx=nx, L=L, n=n)
    for lap in ('cfd4', 'cfd2'):
        out['phi_' + lap] = O.poisson_solve(O.poisson_rhs(n, -1.0, nx), hx, lap)
    np.savez_compressed(os.path.join(HERE, "poisson_32.npz"), **out)


def steps():
    g = O.Grid(16, 32, 4 * np.pi, 8.0, ht=0.05)
    prof = lambda x, v: (1 + 0.5 * np.cos(0.5 * x)) * O.maxwellian(v)
    run = O.OracleRun(g, O.initial_distribution(g, prof))
    out = dict(nx=g.nx, nv=g.nv, L=g.L, vmax=g.vmax, ht=g.ht, f0=run.fc.copy())
    picard, resid, gm = [], [], []
    for s in range(3):
        e = run.step()
        picard.append(len(e['picard']))
        resid.append(e['picard'][-1]['residual'])
        gm.append([p['gmres_its'] for p in e['picard']])
        out['f%d' % (s + 1)] = run.fc.copy()
    out['picard'] = np.array(picard)
    out['residual'] = np.array(resid)
    out['gmres_first'] = np.array([x[0] for x in gm])
    np.savez_compressed(os.path.join(HERE, "landau_steps_16x32_J4.npz"), **out)


if __name__ == "__main__":
    brackets()
    poisson()
    steps()
    print("fixtures written to", HERE)
