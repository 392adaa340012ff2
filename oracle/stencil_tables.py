"""Second, independent transcription of the Arakawa brackets (checker only).

The reference states each bracket twice: matrix-free (``poisson_bracket_array``)
and as the stencil table its ``jacobian()`` assembles into a PETSc matrix.
This module transcribes the tables -- PoissonBracket.pyx:168-180 (J1),
289-301 (J2), 400-402 (J4 = 2 J1 - J2) -- into a dense matrix M_J(H) with
M_J(H) x = {x, H} (times ``factor``).  Tests compare it column by column
with ``vlasov_oracle.bracket`` so a slip in either transcription shows up.

Index convention: unknown (i, j) -> i + nx*j (PETSc natural order).
"""

import numpy as np

STENCIL = 2


def _h(H, i, j):
    nx = H.shape[0]
    return H[i % nx, j]


def j1_entries(H, i, j, fac):
    """PoissonBracket.pyx:168-180: list of ((i', j'), value) for row (i, j)."""
    h = lambda di, dj: _h(H, i + di, j + dj)
    return [
        ((i - 1, j - 1), - (h(-1, 0) - h(0, -1)) * fac),
        ((i - 1, j), - (h(0, +1) - h(0, -1)) * fac - (h(-1, +1) - h(-1, -1)) * fac),
        ((i - 1, j + 1), - (h(0, +1) - h(-1, 0)) * fac),
        ((i, j - 1), + (h(+1, 0) - h(-1, 0)) * fac + (h(+1, -1) - h(-1, -1)) * fac),
        ((i, j + 1), - (h(+1, 0) - h(-1, 0)) * fac - (h(+1, +1) - h(-1, +1)) * fac),
        ((i + 1, j - 1), + (h(+1, 0) - h(0, -1)) * fac),
        ((i + 1, j), + (h(0, +1) - h(0, -1)) * fac + (h(+1, +1) - h(+1, -1)) * fac),
        ((i + 1, j + 1), + (h(0, +1) - h(+1, 0)) * fac),
    ]


def j2_entries(H, i, j, fac):
    """PoissonBracket.pyx:289-301."""
    h = lambda di, dj: _h(H, i + di, j + dj)
    return [
        ((i - 2, j), - (h(-1, +1) - h(-1, -1)) * fac),
        ((i - 1, j - 1), - (h(-2, 0) - h(0, -2)) * fac - (h(-1, +1) - h(+1, -1)) * fac),
        ((i - 1, j + 1), - (h(0, +2) - h(-2, 0)) * fac - (h(+1, +1) - h(-1, -1)) * fac),
        ((i, j - 2), + (h(+1, -1) - h(-1, -1)) * fac),
        ((i, j + 2), - (h(+1, +1) - h(-1, +1)) * fac),
        ((i + 1, j - 1), + (h(+2, 0) - h(0, -2)) * fac + (h(+1, +1) - h(-1, -1)) * fac),
        ((i + 1, j + 1), + (h(0, +2) - h(+2, 0)) * fac + (h(-1, +1) - h(+1, -1)) * fac),
        ((i + 2, j), + (h(+1, +1) - h(+1, -1)) * fac),
    ]


def bracket_matrix(kind, H, grid, factor=1.0):
    """Dense M with (M x)[row] = factor * {x, H}[row] on interior rows."""
    nx, nv = grid.nx, grid.nv
    N = nx * nv
    M = np.zeros((N, N))
    f1 = factor * grid.hx_inv * grid.hv_inv / 12.
    f2 = factor * grid.hx_inv * grid.hv_inv / 24.
    if kind == 'arakawaj1':
        parts = [(j1_entries, f1)]
    elif kind == 'arakawaj2':
        parts = [(j2_entries, f2)]
    elif kind == 'arakawaj4':
        parts = [(j1_entries, +2. * f1), (j2_entries, -1. * f2)]     # :400-402
    else:
        raise ValueError(kind)
    for i in range(nx):
        for j in range(STENCIL, nv - STENCIL):
            row = i + nx * j
            for fn, fac in parts:
                for (ci, cj), val in fn(H, i, j, fac):
                    M[row, (ci % nx) + nx * cj] += val
    return M
