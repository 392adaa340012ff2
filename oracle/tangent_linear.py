"""Tangent-linear model of the implicit timestep (checker only).

TEST INFRASTRUCTURE: only tests/ may use this module, as the CPU FP64
reference of the GPU adjoint.

A converged timestep f0 -> f1 satisfies G(f1, f0) = 0 with
    G(f1, f0) = (f1 - f0)/dt + R B((f0 + f1)/2, H),
    H = h0 + h1[f1]/2 + h1[f0]/2          (h2 = 0: no external field),
where R keeps rows 2..nv-3 and h1[f] is the mean-free potential of the
density of f (Toolbox.pyx:38-98, PETScPoissonSolver.pyx:58-74) -- exactly
the functions ``residual``, ``h_average``, ``density``, ``poisson_rhs``,
``poisson_solve`` and ``potential_to_hamiltonian`` of vlasov_oracle.
Differentiating (B is bilinear, h1 linear):
    A  = dG/df1 =  I/dt + 1/2 R B(., H) + 1/2 R B(fbar, h1[.]),
    A0 = dG/df0 = -I/dt + 1/2 R B(., H) + 1/2 R B(fbar, h1[.]),
and the tangent-linear step is d1 = -A^-1 A0 d0.  Here A and A0 are
assembled densely column by column from those oracle functions and solved
with numpy.linalg.solve, so nothing of the GPU's transposition or Krylov
path is shared.  Pinned in tests/test_oracle_tangent.py: A u equals the
central difference of G (exact, G is quadratic in f1), and the assembled
matrices reproduce the matrix-free application.
"""

import numpy as np

from . import vlasov_oracle as O


def h1_of(f, grid, charge, laplace='cfd4'):
    """h1[f]: density -> Poisson right-hand side -> mean-free phi -> h1."""
    n = O.density(f, grid)
    phi = O.poisson_solve(O.poisson_rhs(n, charge, grid.nx), grid.hx, laplace)
    return O.potential_to_hamiltonian(phi, grid.nx)


def step_residual(kind, f1, f0, h0, grid, charge, laplace='cfd4'):
    """G(f1, f0) with H built from f1 and f0 (the converged Picard fixed point)."""
    H = O.h_average(h0, h1_of(f1, grid, charge, laplace), h1_of(f0, grid, charge, laplace),
                    np.zeros(grid.nx), np.zeros(grid.nx))
    return O.residual(kind, f1, f0, H, grid)


def linear_apply(kind, u, f0, f1, h0, grid, charge, which='A', laplace='cfd4'):
    """A u (which='A') or A0 u (which='A0') about the step (f0, f1)."""
    H = O.h_average(h0, h1_of(f1, grid, charge, laplace), h1_of(f0, grid, charge, laplace),
                    np.zeros(grid.nx), np.zeros(grid.nx))
    fbar = 0.5 * f0 + 0.5 * f1
    sign = 1.0 if which == 'A' else -1.0
    y = sign * grid.ht_inv * u
    O.bracket_add(kind, u, H, y, 0.5, grid)
    hu = np.broadcast_to(h1_of(u, grid, charge, laplace)[:, None], u.shape).copy()
    O.bracket_add(kind, fbar, hu, y, 0.5, grid)
    return y


def dense(kind, f0, f1, h0, grid, charge, which='A', laplace='cfd4'):
    """The matrix of A or A0 (natural order of the oracle's (nx, nv) ravel)."""
    N = grid.nx * grid.nv
    M = np.empty((N, N))
    e = np.zeros(N)
    for k in range(N):
        e[k] = 1.0
        M[:, k] = linear_apply(kind, e.reshape(grid.nx, grid.nv), f0, f1, h0, grid, charge, which,
                               laplace).ravel()
        e[k] = 0.0
    return M


def tangent_linear(kind, states, d0, h0, grid, charge, laplace='cfd4'):
    """Propagate a perturbation d0 of states[0] through the steps
    states[n] -> states[n+1]: returns d_K."""
    d = np.asarray(d0, dtype=np.float64).ravel()
    for n in range(len(states) - 1):
        A = dense(kind, states[n], states[n + 1], h0, grid, charge, 'A', laplace)
        A0 = dense(kind, states[n], states[n + 1], h0, grid, charge, 'A0', laplace)
        d = -np.linalg.solve(A, A0 @ d)
    return d.reshape(grid.nx, grid.nv)


def field_energy(f, grid, charge, laplace='cfd4'):
    """E = 1/2 q sum_i (phi_{i+1} - phi_i)^2 / hx of the potential of f
    (potential.pyx:48-59, calculate_energy, with its charge factor)."""
    phi = O.poisson_solve(O.poisson_rhs(O.density(f, grid), charge, grid.nx), grid.hx, laplace)
    d = np.roll(phi, -1) - phi
    return 0.5 * charge * (d * d).sum() / grid.hx


def field_energy_gradient(f, grid, charge, laplace='cfd4'):
    """dE/df, by the chain rule through density -> rhs -> phi (all linear)."""
    phi = O.poisson_solve(O.poisson_rhs(O.density(f, grid), charge, grid.nx), grid.hx, laplace)
    g = charge * ((phi - np.roll(phi, 1)) - (np.roll(phi, -1) - phi)) / grid.hx      # dE/dphi
    # phi = S b with S the (symmetric) mean-free solve, b = -q (n - mean n), n = hv sum_j f
    w = O.poisson_solve(g, grid.hx, laplace)                                 # S^T g
    dn = -charge * (w - w.mean())
    return np.broadcast_to((grid.hv * dn)[:, None], f.shape).copy()


def tangent_linear_mf(kind, states, d0, h0, grid, charge, laplace='cfd4', rtol=1e-13):
    """The tangent-linear propagation of tangent_linear() without assembling
    A: each step solves A d1 = -A0 d0 by SciPy's GMRES on the matrix-free
    linear_apply, preconditioned by the oracle's kinetic P^-1 (SciPy band
    solves), to rtol.  For grids where the dense matrices do not fit (256 x 512
    and up); pinned against the dense model on small grids."""
    from scipy.sparse.linalg import LinearOperator, gmres
    N = grid.nx * grid.nv
    shape = (grid.nx, grid.nv)
    pc = O.Preconditioner(grid, 'scipy')
    d = np.asarray(d0, dtype=np.float64).reshape(shape)
    for n in range(len(states) - 1):
        f0, f1 = states[n], states[n + 1]
        A = LinearOperator((N, N), matvec=lambda u: linear_apply(kind, u.reshape(shape), f0, f1, h0, grid, charge,
                                                                 'A', laplace).ravel(), dtype=np.float64)
        M = LinearOperator((N, N), matvec=lambda u: pc.apply(u.reshape(shape)).ravel(), dtype=np.float64)
        b = -linear_apply(kind, d, f0, f1, h0, grid, charge, 'A0', laplace).ravel()
        x, info = gmres(A, b, M=M, rtol=rtol, atol=0.0, restart=60, maxiter=50)
        if info != 0:
            raise RuntimeError("tangent-linear GMRES did not converge (info %d)" % info)
        d = x.reshape(shape)
    return d
