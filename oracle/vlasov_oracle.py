"""NumPy/SciPy restatement of the viVlasov1D implicit timestep (checker only).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may use this module, and only as the checker.

Array convention: a 2D field is indexed ``f[i, j]`` with ``i`` the periodic x
index and ``j`` the v index, shape ``(nx, nv)``.  That is the reference's own
view of a PETSc DMDA vector (``vlasov/toolbox/VIDA.pyx:12-51``: a Fortran-order
reshape of natural order ``i + nx*j``).  The C-contiguous ``(nv, nx)`` array
``f.T`` holds the same bytes in the same order; the product library works on
that natural order directly.

How each function is pinned (the tests in ``tests/test_oracle_*.py``):

* ``bracket_*`` (J1/J2/J4) -- (i) the matrix-free stencil transcribed
  term by term; (ii) cross-checked against ``stencil_tables``, an
  independent transcription of the reference's assembled ``jacobian()``
  stencil tables; (iii) Arakawa identities: antisymmetry B(f,h) = -B(h,f)
  pointwise and sum B = sum f B = sum h B = 0 on fields vanishing near the
  v-boundaries; (iv) closed-form Fourier symbols for separable inputs.
* ``residual`` / ``jacobian`` -- closed form; F is affine in f for fixed H,
  so central differences of F equal J e_k (brute force on tiny grids).
* ``Preconditioner`` -- LAPACK ``zgbtrf/zgbtrs`` exactly as the reference
  calls them, cross-checked against the reference's second statement
  (numpy rfft + ``scipy.linalg.solve_banded``), against a dense assembled P
  (P (P^-1 x) = x) and against the Fourier symbol of 1/2 B_J1(., h0).
* ``poisson_*`` -- exact circulant eigenvalues vs a dense least-squares
  solve of the assembled stencil matrix; residual of the solve.
* ``density`` / invariants -- closed-form Maxwellian moments.
* ``gmres`` -- dense ``numpy.linalg.solve`` on small systems, restart
  equivalence, monotone residual history.
* ``collision_add`` / ``collision_factor`` -- (i) against ``collision_stencil``,
  a transcription of the reference's assembled ``jacobian()`` of the same
  operators; (ii) exact discrete conservation of mass, momentum and energy
  when the moments are those of f and f vanishes near the v-boundaries;
  (iii) a drifting Maxwellian lies in the null space to O(hv^2) (halving hv
  quarters the remainder); the diffusion part alone has sum f C(f) > 0;
  a = 1/T for a Maxwellian.
* ``double_bracket_*`` -- the Arakawa trilinear antisymmetry gives
  sum H D_ffh = nu sum {F,H}^2, sum F D_ffh = 0, sum F D_fhh = nu sum {F,H}^2,
  sum H D_fhh = 0; the Jacobian terms equal central differences of the
  function terms (quadratic / linear in f, so exact up to rounding).
* ``OracleRun`` (Picard driver) -- each stop / revert / abort branch forced
  from the host; invariant drift bounds.
"""

import math

import numpy as np
from scipy.linalg import lapack, solve_banded

STENCIL = 2   # vlasov/run/run_base.py:48-49


# --------------------------------------------------------------------------
# grid                                                    run_base.py:118-137
# --------------------------------------------------------------------------
class Grid:
    """Grid constants as ``vlasov/core/Grid.pyx:14-51`` stores them.

    x_i = i*hx (periodic DMDA uniform coordinates), v_j = vmin + j*hv
    (non-periodic), hx = L/nx, hv = (vmax - vmin)/(nv - 1)
    (``run_base.py:130-137, 252-254``).
    """

    def __init__(self, nx, nv, L, vmax, vmin=None, ht=0.1, nt=1):
        if vmin is None:
            vmin = -vmax
        self.nx, self.nv, self.nt = int(nx), int(nv), int(nt)
        self.L, self.vmin, self.vmax = float(L), float(vmin), float(vmax)
        self.ht = float(ht)
        self.hx = self.L / self.nx
        self.hv = (self.vmax - self.vmin) / (self.nv - 1)
        self.stencil = STENCIL
        self.x = np.array([0.0 + i * self.hx for i in range(self.nx)])
        self.v = np.array([self.vmin + j * self.hv for j in range(self.nv)])
        self.v2 = np.power(self.v, 2)
        self.ht_inv = 1. / self.ht
        self.hx_inv = 1. / self.hx
        self.hv_inv = 1. / self.hv
        self.hx2 = self.hx ** 2
        self.hv2 = self.hv ** 2
        self.hx2_inv = 1. / self.hx2
        self.hv2_inv = 1. / self.hv2


# --------------------------------------------------------------------------
# Poisson brackets                        components/PoissonBracket.pyx
# --------------------------------------------------------------------------
def _s(a, di, dj):
    """a[(i+di) mod nx, j+dj] for all i and the interior rows 2 <= j < nv-2."""
    nv = a.shape[1]
    return np.roll(a, -di, axis=0)[:, STENCIL + dj: nv - STENCIL + dj]


def _j1_terms(x, h):
    """jpp + jpc + jcp of PoissonBracket.pyx:100-111 (left-to-right order)."""
    S = _s
    jpp = (S(x, 1, 0) - S(x, -1, 0)) * (S(h, 0, 1) - S(h, 0, -1)) \
        - (S(x, 0, 1) - S(x, 0, -1)) * (S(h, 1, 0) - S(h, -1, 0))
    jpc = S(x, 1, 0) * (S(h, 1, 1) - S(h, 1, -1)) \
        - S(x, -1, 0) * (S(h, -1, 1) - S(h, -1, -1)) \
        - S(x, 0, 1) * (S(h, 1, 1) - S(h, -1, 1)) \
        + S(x, 0, -1) * (S(h, 1, -1) - S(h, -1, -1))
    jcp = S(x, 1, 1) * (S(h, 0, 1) - S(h, 1, 0)) \
        - S(x, -1, -1) * (S(h, -1, 0) - S(h, 0, -1)) \
        - S(x, -1, 1) * (S(h, 0, 1) - S(h, -1, 0)) \
        + S(x, 1, -1) * (S(h, 1, 0) - S(h, 0, -1))
    return jpp + jpc + jcp


def _j2_terms(x, h):
    """jcc + jpc + jcp of PoissonBracket.pyx:221-232."""
    S = _s
    jcc = (S(x, 1, 1) - S(x, -1, -1)) * (S(h, -1, 1) - S(h, 1, -1)) \
        - (S(x, -1, 1) - S(x, 1, -1)) * (S(h, 1, 1) - S(h, -1, -1))
    jpc = S(x, 2, 0) * (S(h, 1, 1) - S(h, 1, -1)) \
        - S(x, -2, 0) * (S(h, -1, 1) - S(h, -1, -1)) \
        - S(x, 0, 2) * (S(h, 1, 1) - S(h, -1, 1)) \
        + S(x, 0, -2) * (S(h, 1, -1) - S(h, -1, -1))
    jcp = S(x, 1, 1) * (S(h, 0, 2) - S(h, 2, 0)) \
        - S(x, -1, -1) * (S(h, -2, 0) - S(h, 0, -2)) \
        - S(x, -1, 1) * (S(h, 0, 2) - S(h, -2, 0)) \
        + S(x, 1, -1) * (S(h, 2, 0) - S(h, 0, -2))
    return jcc + jpc + jcp


def bracket_add(kind, x, h, y, factor, grid):
    """``y += factor * {x, h}`` on interior rows, in place.

    J1: ``y += factor*(jpp+jpc+jcp)*af1``      (PoissonBracket.pyx:80-113)
    J2: ``y += factor*(jcc+jpc+jcp)*af2``      (PoissonBracket.pyx:200-234)
    J4: ``y += factor*(2*rJ1 - rJ2)``          (PoissonBracket.pyx:335-389)
    with af1 = hx_inv*hv_inv/12, af2 = hx_inv*hv_inv/24.
    Rows j < 2 and j >= nv-2 receive nothing (:94, :215, :354).
    """
    af1 = grid.hx_inv * grid.hv_inv / 12.
    af2 = grid.hx_inv * grid.hv_inv / 24.
    inner = y[:, STENCIL:grid.nv - STENCIL]
    if kind == 'arakawaj1':
        inner += factor * _j1_terms(x, h) * af1
    elif kind == 'arakawaj2':
        inner += factor * _j2_terms(x, h) * af2
    elif kind == 'arakawaj4':
        r1 = _j1_terms(x, h) * af1
        r2 = _j2_terms(x, h) * af2
        inner += factor * (2. * r1 - r2)
    else:
        raise ValueError("unknown Poisson bracket %r" % kind)
    return y


def bracket(kind, x, h, grid, factor=1.0):
    y = np.zeros_like(x)
    return bracket_add(kind, x, h, y, factor, grid)


# --------------------------------------------------------------------------
# Hamiltonian composition                   PETScVlasovSolver.pyx:142-157
# --------------------------------------------------------------------------
def h_average(h0, h1p, h1h, h2p, h2h):
    """H_ave = H0; += .5 H1p; += .5 H1h; += .5 H2p; += .5 H2h.

    ``h0`` is the v-only kinetic part (Toolbox.pyx:187) of shape (nv,), the
    others are x-only (Toolbox.pyx:59) of shape (nx,); the result is the 2D
    (nx, nv) field with the reference's addition order.
    """
    H = np.broadcast_to(h0[None, :], (h1p.shape[0], h0.shape[0])).copy()
    H = H + .5 * h1p[:, None]
    H = H + .5 * h1h[:, None]
    H = H + .5 * h2p[:, None]
    H = H + .5 * h2h[:, None]
    return H


def kinetic_hamiltonian(grid, mass=1.0):
    """h0_j = 0.5 * v_j**2 * mass  (Toolbox.pyx:175-187)."""
    return 0.5 * grid.v ** 2 * mass


# --------------------------------------------------------------------------
# residual and Jacobian                        PETScNLVlasovMP.pyx:33-58
# --------------------------------------------------------------------------
def residual(kind, f, fh, H, grid):
    """function_solver (PETScNLVlasovMP.pyx:43-58), 'point' time derivative.

    Fave = (0 + .5 Fh) + .5 F; Fder = (0 + F) - Fh; Y = 0 + Fder/dt;
    Y += B(Fave, H, 1.0).  TimeDerivative.function is Y.axpy(ht_inv, F)
    (TimeDerivative.pyx:56-57).
    """
    fave = 0.5 * fh + 0.5 * f
    fder = f - fh
    y = grid.ht_inv * fder
    return bracket_add(kind, fave, H, y, 1.0, grid)


def jacobian(kind, u, H, grid):
    """jacobian_solver (PETScNLVlasovMP.pyx:33-40): Y = u/dt + .5 B(u, H)."""
    y = grid.ht_inv * u
    return bracket_add(kind, u, H, y, 0.5, grid)


# --------------------------------------------------------------------------
# dissipation                 components/CollisionOperator.pyx:68-250
#                             components/DoubleBracket.pyx:67-97
#                             toolbox/Toolbox.pyx:153-165
# --------------------------------------------------------------------------
def collision_factor(n, u, e):
    """a_i = n_i n_i / (n_i e_i - u_i u_i)  (Toolbox.pyx:153-165); u, e are the
    raw first and second velocity moments."""
    return n * n / (n * e - u * u)


def collision_add(kind, f, y, n, U, a, factor, grid, coll_freq, coll_diff=1., coll_drag=1.):
    """``y += C(f)`` on rows 2..nv-3, in place (CollisionOperatorT.function,
    CollisionOperator.pyx:72-110; CollisionOperatorE.function, :164-202).

    drag_fac = -factor*nu*drag*hv_inv*0.5, diff_fac = -factor*nu*diff*hv2_inv
    (left to right), u = U/n, and with f+, f0, f- the rows j+1, j, j-1:
      collt: drag = ((v+ - u) f+ - (v- - u) f-) * a;  diff = (f+ - 2 f0) + f-
      colle: drag = (v+ - u) f+ - (v- - u) f-;        diff = ((f+ - 2 f0) + f-) / a
      y += drag_fac*drag + diff_fac*diff
    Nothing is added when nu == 0 or factor == 0 (:87).
    """
    if coll_freq == 0. or factor == 0.:
        return y
    drag_fac = -factor * coll_freq * coll_drag * grid.hv_inv * 0.5
    diff_fac = -factor * coll_freq * coll_diff * grid.hv2_inv
    s, nv = STENCIL, grid.nv
    u = (U / n)[:, None]
    aa = a[:, None]
    vp = grid.v[s + 1:nv - s + 1][None, :]
    vm = grid.v[s - 1:nv - s - 1][None, :]
    fp, f0, fm = f[:, s + 1:nv - s + 1], f[:, s:nv - s], f[:, s - 1:nv - s - 1]
    if kind == 'collt':
        drag = ((vp - u) * fp - (vm - u) * fm) * aa
        diff = (fp - 2. * f0) + fm
    elif kind == 'colle':
        drag = (vp - u) * fp - (vm - u) * fm
        diff = ((fp - 2. * f0) + fm) / aa
    else:
        raise ValueError("unknown collision operator %r" % kind)
    y[:, s:nv - s] += drag_fac * drag + diff_fac * diff
    return y


def collision_stencil(kind, n, U, a, factor, grid, coll_freq, coll_diff=1., coll_drag=1.):
    """Second statement of the collision operators: the (j-1, j, j+1) entries
    the reference's assembled ``jacobian()`` adds for row j
    (CollisionOperatorT.jacobian :147-153; CollisionOperatorE.jacobian
    :240-246), as three (nx, nv) arrays (zero outside rows 2..nv-3)."""
    drag_fac = -factor * coll_freq * coll_drag * grid.hv_inv * 0.5
    diff_fac = -factor * coll_freq * coll_diff * grid.hv2_inv
    nx, nv, s = grid.nx, grid.nv, STENCIL
    lo, mid, hi = np.zeros((nx, nv)), np.zeros((nx, nv)), np.zeros((nx, nv))
    u = U / n
    for j in range(s, nv - s):
        if kind == 'collt':
            lo[:, j] = -drag_fac * (grid.v[j - 1] - u) * a + diff_fac
            mid[:, j] = -2. * diff_fac
            hi[:, j] = +drag_fac * (grid.v[j + 1] - u) * a + diff_fac
        else:
            lo[:, j] = -drag_fac * (grid.v[j - 1] - u) + diff_fac / a
            mid[:, j] = -2. * diff_fac / a
            hi[:, j] = +drag_fac * (grid.v[j + 1] - u) + diff_fac / a
    return lo, mid, hi


def double_bracket_function(kind, bkind, fave, H, y, factor, grid, coll_freq):
    """``y += D(Fave)`` in place, with b = 0 + {Fave, H} (factor 1.0):
      ffh: y += (-nu*factor) {Fave, b}   (DoubleBracketFFH.function, DoubleBracket.pyx:79-82)
      fhh: y += (-nu*factor) {b, H}      (DoubleBracketFHH.function, :94-97)
    ``bkind`` is the Poisson bracket (J1/J2/J4) the solver was built with."""
    b = bracket(bkind, fave, H, grid)
    if kind == 'ffh':
        return bracket_add(bkind, fave, b, y, -coll_freq * factor, grid)
    if kind == 'fhh':
        return bracket_add(bkind, b, H, y, -coll_freq * factor, grid)
    raise ValueError("unknown double bracket %r" % kind)


def double_bracket_jacobian(kind, bkind, u, fave, H, y, factor, grid, coll_freq):
    """``y += D'(Fave) u`` in place:
      ffh: b = {Fave, H}; y += (-nu f) {u, b}; b = {u, H}; y += (-nu f) {Fave, b}
           (DoubleBracketFFH.jacobian, DoubleBracket.pyx:69-76)
      fhh: b = {u, H}; y += (-nu f) {b, H}   (DoubleBracketFHH.jacobian, :88-91)"""
    if kind == 'ffh':
        b = bracket(bkind, fave, H, grid)
        bracket_add(bkind, u, b, y, -coll_freq * factor, grid)
        b = bracket(bkind, u, H, grid)
        return bracket_add(bkind, fave, b, y, -coll_freq * factor, grid)
    if kind == 'fhh':
        b = bracket(bkind, u, H, grid)
        return bracket_add(bkind, b, H, y, -coll_freq * factor, grid)
    raise ValueError("unknown double bracket %r" % kind)


class Dissipation:
    """The dissipation term of the solver (PETScVlasovSolver.pyx:101-102,
    chosen by config.py:148-163): a collision operator ('collt' | 'colle',
    needing the moments set by set_moments, PETScVlasovSolver.pyx:123-132) or
    a double bracket ('ffh' | 'fhh').  ``mom`` holds the x-vectors
    np, up, ap (current) and nh, uh, ah (history): density, raw velocity
    density and collision factor."""

    def __init__(self, kind, coll_freq, coll_diff=1., coll_drag=1.):
        if kind not in ('collt', 'colle', 'ffh', 'fhh'):
            raise ValueError("unknown dissipation %r" % kind)
        self.kind = kind
        self.coll_freq, self.coll_diff, self.coll_drag = float(coll_freq), float(coll_diff), float(coll_drag)

    @property
    def collisions(self):
        return self.kind in ('collt', 'colle')

    def _coll(self, f, y, n, U, a, factor, grid):
        return collision_add(self.kind, f, y, n, U, a, factor, grid, self.coll_freq, self.coll_diff, self.coll_drag)

    def function(self, bkind, f, fh, fave, H, y, grid, mom):
        """The terms function_solver adds after B(Fave, H) (PETScNLVlasovMP.pyx:56-58)."""
        if self.collisions:
            self._coll(f, y, mom['np'], mom['up'], mom['ap'], 0.5, grid)
            self._coll(fh, y, mom['nh'], mom['uh'], mom['ah'], 0.5, grid)
            return y
        return double_bracket_function(self.kind, bkind, fave, H, y, 1.0, grid, self.coll_freq)

    def jacobian(self, bkind, u, fave, H, y, grid, mom):
        """The terms jacobian_solver adds after .5 B(u, H) (PETScNLVlasovMP.pyx:38-39)."""
        if self.collisions:
            return self._coll(u, y, mom['np'], mom['up'], mom['ap'], 0.5, grid)
        return double_bracket_jacobian(self.kind, bkind, u, fave, H, y, 0.5, grid, self.coll_freq)


# --------------------------------------------------------------------------
# explicit initial guess        run_base.py:786-927,
#                               solvers/explicit/PETScArakawaSymplectic.pyx:62-97,
#                               solvers/explicit/PETScArakawaRungeKutta.pyx:20-98
# The reference's steppers cannot run as written (bracket arguments passed in
# the wrong order, an accumulator never cleared, an undefined rk18_J4); these
# are the schemes they state, with the time derivative of the implicit
# residual, df/dt = -{f, H}, and each tendency formed afresh.
# --------------------------------------------------------------------------
def tendency(kind, x, h, grid):
    """-{x, h} on rows 2..nv-3 (0 + (-1.0) * bracket), zero elsewhere."""
    return bracket(kind, x, h, grid, factor=-1.0)


def symplectic_kinetic(kind, x, h0, factor, grid, niter=1):
    """PETScArakawaSymplectic.kinetic (:62-77): x += factor*ht/niter * (-{x, H0})."""
    return x + factor * grid.ht / float(niter) * tendency(kind, x, h0, grid)


def symplectic_potential(kind, x, h12, factor, grid, niter=1):
    """PETScArakawaSymplectic.potential (:80-97): x += factor*ht/niter * (-{x, H1 + H2})."""
    return x + factor * grid.ht / float(niter) * tendency(kind, x, h12, grid)


def symplectic4_coefficients():
    """run_base.py:892-901 (fourth-order Forest-Ruth / Yoshida composition)."""
    fac2 = 2. ** (1. / 3.)
    c1 = 0.5 / (2. - fac2)
    c2 = c1 * (1. - fac2)
    d1 = 1. / (2. - fac2)
    d2 = -d1 * fac2
    return (c1, c2, c2, c1), (d1, d2, d1)


def rk4_16(kind, x, h, grid, niter=1):
    """PETScArakawaRungeKutta.rk4_16 (:20-57), classical RK4, h = h0 + h1 + h2."""
    s = grid.ht / float(niter)
    t1 = tendency(kind, x, h, grid)
    t2 = tendency(kind, x + 0.5 * grid.ht / float(niter) * t1, h, grid)
    t3 = tendency(kind, x + 0.5 * grid.ht / float(niter) * t2, h, grid)
    t4 = tendency(kind, x + 1.0 * grid.ht / float(niter) * t3, h, grid)
    return x + s * (t1 + 2. * t2 + 2. * t3 + t4) / 6.


def rk4_18(kind, x, h, grid, niter=1):
    """PETScArakawaRungeKutta.rk4_18 (:61-98), the 3/8 rule, h = h0 + h1 + h2."""
    s = grid.ht / float(niter)
    t1 = tendency(kind, x, h, grid)
    t2 = tendency(kind, x + 1. / 3. * grid.ht / float(niter) * t1, h, grid)
    t3 = tendency(kind, x - 1. / 3. * grid.ht / float(niter) * t1 + s * t2, h, grid)
    t4 = tendency(kind, x + s * t1 - s * t2 + s * t3, h, grid)
    return x + s * (t1 + 3. * t2 + 3. * t3 + t4) / 8.


GUESSES = ('symplectic2', 'symplectic4', 'rk4', 'rk4_16', 'rk4_18')


def residual_diss(kind, f, fh, H, grid, diss, mom=None):
    """function_solver with its dissipation term (PETScNLVlasovMP.pyx:43-58)."""
    y = residual(kind, f, fh, H, grid)
    return diss.function(kind, f, fh, 0.5 * fh + 0.5 * f, H, y, grid, mom)


def jacobian_diss(kind, u, fave, H, grid, diss, mom=None):
    """jacobian_solver with its dissipation term (PETScNLVlasovMP.pyx:33-40);
    ``fave`` is the average the last function_solver call formed."""
    y = jacobian(kind, u, H, grid)
    return diss.jacobian(kind, u, fave, H, y, grid, mom)


# --------------------------------------------------------------------------
# kinetic tensor-product preconditioner
#   preconditioner/TensorProduct.pyx:109-116 (eigenvalues)
#   preconditioner/TensorProductKineticFast.pyx:61-90, 102-272
#   preconditioner/TensorProductKineticSciPy.pyx:53-137 (second statement)
# --------------------------------------------------------------------------
def kinetic_eigenvalues(n):
    """TensorProduct.compute_eigenvalues (TensorProduct.pyx:109-116)."""
    eigen = np.empty(n, dtype=np.complex128)
    for i in range(n):
        eigen[i] = np.exp(2.j * np.pi * float(i) / n * (n - 1)) \
            - np.exp(2.j * np.pi * float(i) / n)
    return np.fft.ifftshift(eigen)


def kinetic_bands(grid, eigen_k):
    """Sub-, main and super-diagonal of T_k (KineticFast.pyx:248-260).

    diagm_j = e*0.5*(2 hv v_j - hv^2)*a, diag_j = e*1.0*(4 hv v_j)*a + 1/dt,
    diagp_j = e*0.5*(2 hv v_j + hv^2)*a with a = 0.5*hx_inv*hv_inv/12.
    Complex-times-real is evaluated componentwise, as C does it.
    """
    a = 0.5 * grid.hx_inv * grid.hv_inv / 12.
    er, ei = float(eigen_k.real), float(eigen_k.imag)
    gm = 2. * grid.hv * grid.v - grid.hv2
    g0 = 4. * grid.hv * grid.v
    gp = 2. * grid.hv * grid.v + grid.hv2
    diagm = ((er * 0.5) * gm * a) + 1j * ((ei * 0.5) * gm * a)
    diag = ((er * 1.0) * g0 * a + grid.ht_inv) + 1j * ((ei * 1.0) * g0 * a)
    diagp = ((er * 0.5) * gp * a) + 1j * ((ei * 0.5) * gp * a)
    return diagm, diag, diagp


class IdentityPreconditioner:
    """No preconditioner (preconditioner_type None, TensorProduct.pyx:66-86
    returns None): function_snes_mult / mult apply F and J directly."""

    def apply(self, x):
        return x


class Preconditioner:
    """P^-1 x = irfft_x( T_k^-1 rfft_x(x) )   (TensorProductKinetic.pyx:146-160).

    ``variant='lapack'`` follows TensorProductKineticFast: LAPACK band
    storage with LDA=4, KL=KU=1 (:61-68, :270-272), zgbtrf once (:88-90),
    zgbtrs per mode and apply (:223-224).  ``variant='scipy'`` follows
    TensorProductKineticSciPy: 3-row band storage and solve_banded
    (:115-137, :106-108).  Both use the backward-normalised real FFT pair
    (pyFFTW normalises the inverse exactly as numpy.fft.irfft).
    """

    def __init__(self, grid, variant='lapack'):
        self.grid = grid
        self.variant = variant
        nx, nv = grid.nx, grid.nv
        self.eigen = kinetic_eigenvalues(nx)
        self.nmodes = nx // 2 + 1
        self.bands = [kinetic_bands(grid, self.eigen[k]) for k in range(self.nmodes)]
        self.info = []
        if variant == 'lapack':
            self.lu = []
            for k in range(self.nmodes):
                dm, d, dp = self.bands[k]
                ab = np.zeros((4, nv), dtype=np.complex128, order='F')
                ab[1, 1:] = dp[:-1]
                ab[2, :] = d
                ab[3, :-1] = dm[1:]
                lu, piv, info = lapack.zgbtrf(ab, 1, 1)
                if info != 0:
                    # KineticFast.pyx:89-90 prints and continues
                    print("   ERROR in LU Decomposition.")
                self.info.append(info)
                self.lu.append((lu, piv))
        elif variant == 'scipy':
            self.ab3 = []
            for k in range(self.nmodes):
                dm, d, dp = self.bands[k]
                ab = np.zeros((3, nv), dtype=np.complex128)
                ab[0, 1:] = dp[:-1]
                ab[1, :] = d
                ab[2, :-1] = dm[1:]
                self.ab3.append(ab)
        else:
            raise ValueError(variant)

    def pivots_are_identity(self):
        if self.variant != 'lapack':
            return True
        # scipy reports LAPACK's ipiv 0-based: row j pivots with itself <=> ipiv[j] == j
        return all(np.array_equal(piv, np.arange(self.grid.nv)) for (_, piv) in self.lu)

    def solve_modes(self, X):
        """Solve T_k c = b for each mode row X[k, :] (shape (nmodes, nv))."""
        out = np.empty_like(X)
        for k in range(self.nmodes):
            if self.variant == 'lapack':
                lu, piv = self.lu[k]
                sol, info = lapack.zgbtrs(lu, 1, 1, X[k, :].copy(), piv)
                out[k, :] = sol
            else:
                out[k, :] = solve_banded((1, 1), self.ab3[k], X[k, :].copy())
        return out

    def apply(self, x):
        X = np.fft.rfft(x, axis=0)              # (nx//2+1, nv)
        C = self.solve_modes(X)
        return np.fft.irfft(C, n=self.grid.nx, axis=0)

    def dense_matrix(self):
        """The operator P (not P^-1), assembled densely for tiny grids."""
        nx, nv = self.grid.nx, self.grid.nv
        N = nx * nv
        P = np.zeros((N, N))
        for col in range(N):
            e = np.zeros(N)
            e[col] = 1.0
            x = e.reshape((nv, nx)).T                # natural order i + nx*j
            X = np.fft.rfft(x, axis=0)
            Y = np.empty_like(X)
            for k in range(self.nmodes):
                dm, d, dp = self.bands[k]
                T = np.diag(d) + np.diag(dp[:-1], 1) + np.diag(dm[1:], -1)
                Y[k, :] = T @ X[k, :]
            y = np.fft.irfft(Y, n=nx, axis=0)
            P[:, col] = y.T.reshape(N)
        return P


def potential_bands(grid, phi, eigen_l):
    """Sub-, main and super-diagonal of T_l along x for v-mode l
    (TensorProductPotentialFast.pyx:229-264; PotentialSciPy.pyx:126-164):
    diagm_i = e*1.0*(phi_i - phi_{i-1})*a, diag_i = e*2.0*(phi_{i+1} -
    phi_{i-1})*a + 1/dt, diagp_i = e*1.0*(phi_{i+1} - phi_i)*a, indices mod nx,
    a = 0.5*hx_inv*hv_inv/12.  The band matrix takes diagp[:-1] and diagm[1:],
    so the periodic corners are dropped.
    """
    a = 0.5 * grid.hx_inv * grid.hv_inv / 12.
    er, ei = float(eigen_l.real), float(eigen_l.imag)
    phi = np.asarray(phi, dtype=np.float64)
    phim, phip = np.roll(phi, 1), np.roll(phi, -1)
    gm, g0, gp = phi - phim, phip - phim, phip - phi
    diagm = ((er * 1.0) * gm * a) + 1j * ((ei * 1.0) * gm * a)
    diag = ((er * 2.0) * g0 * a + grid.ht_inv) + 1j * ((ei * 2.0) * g0 * a)
    diagp = ((er * 1.0) * gp * a) + 1j * ((ei * 1.0) * gp * a)
    return diagm, diag, diagp


class PotentialPreconditioner:
    """P^-1 x = irfft_v( T_l^-1 rfft_v(x) )  (TensorProductPotential.pyx:157-173):
    a real FFT along v for every x, then for every v-mode l = 0..nv/2 the
    tridiagonal system T_l along x, whose coefficients depend on phi and are
    rebuilt by update() (update_matrices, called from update_history at every
    timestep, PETScVlasovSolver.pyx:135-139).  ``variant`` as in
    Preconditioner: 'lapack' (zgbtrf/zgbtrs, PotentialFast.pyx:230-264) or
    'scipy' (solve_banded, PotentialSciPy.pyx:96-104).
    """

    def __init__(self, grid, phi, variant='lapack'):
        self.grid = grid
        self.variant = variant
        self.eigen = kinetic_eigenvalues(grid.nv)        # compute_eigenvalues(nv)
        self.nmodes = grid.nv // 2 + 1
        self.update(phi)

    def update(self, phi):
        nx = self.grid.nx
        self.bands = [potential_bands(self.grid, phi, self.eigen[l]) for l in range(self.nmodes)]
        self.info = []
        self.lu, self.ab3 = [], []
        for l in range(self.nmodes):
            dm, d, dp = self.bands[l]
            if self.variant == 'lapack':
                ab = np.zeros((4, nx), dtype=np.complex128, order='F')
                ab[1, 1:] = dp[:-1]
                ab[2, :] = d
                ab[3, :-1] = dm[1:]
                lu, piv, info = lapack.zgbtrf(ab, 1, 1)
                if info != 0:
                    print("   ERROR in LU Decomposition.")
                self.info.append(info)
                self.lu.append((lu, piv))
            elif self.variant == 'scipy':
                ab = np.zeros((3, nx), dtype=np.complex128)
                ab[0, 1:] = dp[:-1]
                ab[1, :] = d
                ab[2, :-1] = dm[1:]
                self.ab3.append(ab)
            else:
                raise ValueError(self.variant)

    def pivots_are_identity(self):
        if self.variant != 'lapack':
            return True
        return all(np.array_equal(piv, np.arange(self.grid.nx)) for (_, piv) in self.lu)

    def apply(self, x):
        """x in the oracle layout (nx, nv)."""
        X = np.fft.rfft(x, axis=1)               # (nx, nv//2+1)
        C = np.empty_like(X)
        for l in range(self.nmodes):
            if self.variant == 'lapack':
                lu, piv = self.lu[l]
                C[:, l], info = lapack.zgbtrs(lu, 1, 1, X[:, l].copy(), piv)
            else:
                C[:, l] = solve_banded((1, 1), self.ab3[l], X[:, l].copy())
        return np.fft.irfft(C, n=self.grid.nv, axis=1)

    def dense_matrix(self):
        """The operator P (not P^-1), assembled densely for tiny grids
        (natural order i + nx*j)."""
        nx, nv = self.grid.nx, self.grid.nv
        N = nx * nv
        P = np.zeros((N, N))
        Ts = []
        for l in range(self.nmodes):
            dm, d, dp = self.bands[l]
            Ts.append(np.diag(d) + np.diag(dp[:-1], 1) + np.diag(dm[1:], -1))
        for col in range(N):
            e = np.zeros(N)
            e[col] = 1.0
            x = e.reshape((nv, nx)).T
            X = np.fft.rfft(x, axis=1)
            Y = np.empty_like(X)
            for l in range(self.nmodes):
                Y[:, l] = Ts[l] @ X[:, l]
            P[:, col] = np.fft.irfft(Y, n=nv, axis=1).T.reshape(N)
        return P


# --------------------------------------------------------------------------
# moments and Poisson
#   toolbox/Toolbox.pyx:38-98; poisson/PETScPoissonSolver.pyx:58-74;
#   poisson/PETScPoissonCFD4.pyx:29-100; poisson/PETScPoissonCFD2.pyx:29-100
# --------------------------------------------------------------------------
def density(f, grid):
    """n_i = (sum_j f_ij) * hv, summed sequentially in j (Toolbox.pyx:79-98)."""
    n = np.zeros(f.shape[0])
    for j in range(f.shape[1]):
        n += f[:, j]
    return n * grid.hv


def velocity_density(f, grid):
    """u_i = (sum_j v_j f_ij) * hv, summed in j (Toolbox.pyx:103-123)."""
    u = np.zeros(f.shape[0])
    for j in range(f.shape[1]):
        u += grid.v[j] * f[:, j]
    return u * grid.hv


def energy_density(f, grid):
    """e_i = (sum_j v_j**2 f_ij) * hv, summed in j (Toolbox.pyx:128-148)."""
    e = np.zeros(f.shape[0])
    for j in range(f.shape[1]):
        e += grid.v2[j] * f[:, j]
    return e * grid.hv


def potential_to_hamiltonian(phi, nx):
    """h1_i = phi_i - sum(phi)/nx  (Toolbox.pyx:38-74), kept 1D."""
    return phi - phi.sum() / nx


def poisson_rhs(n, charge, nx):
    """b_i = -(n_i - mean n) * charge  (PETScPoissonSolver.pyx:58-74)."""
    nmean = n.sum() / nx
    return -(n - nmean) * charge


def poisson_matrix_mult(x, hx2_inv, laplace='cfd4'):
    """matrix_mult of PETScPoissonCFD4.pyx:64-78 / PETScPoissonCFD2.pyx:64-78."""
    xm2, xm1 = np.roll(x, 2), np.roll(x, 1)
    xp1, xp2 = np.roll(x, -1), np.roll(x, -2)
    if laplace == 'cfd4':
        return (5. * xm2 - 32. * xm1 + 54. * x - 32. * xp1 + 5. * xp2) * hx2_inv / 12.
    if laplace == 'cfd2':
        return (2. * x - xm1 - xp1) * hx2_inv
    raise ValueError(laplace)


def poisson_function_mult(x, n, charge, hx2_inv, laplace='cfd4'):
    """function_mult: A x + (n - mean n) q  (PETScPoissonCFD4.pyx:83-100)."""
    nmean = n.sum() / n.shape[0]
    return poisson_matrix_mult(x, hx2_inv, laplace) + (n - nmean) * charge


def poisson_eigenvalues(nx, hx, laplace='cfd4'):
    """Symbol of the periodic stencil: lambda_k, k = 0..nx-1 (lambda_0 = 0)."""
    th = 2.0 * np.pi * np.arange(nx) / nx
    if laplace == 'cfd4':
        lam = (54. - 64. * np.cos(th) + 10. * np.cos(2. * th)) / (12. * hx ** 2)
    elif laplace == 'cfd2':
        lam = (2. - 2. * np.cos(th)) / hx ** 2
    else:
        raise ValueError(laplace)
    lam[0] = 0.0
    return lam


def poisson_solve(b, hx, laplace='cfd4'):
    """Mean-free solution of A phi = b on the periodic grid.

    The reference solves the singular system with CG + LU and a constant
    null space to rtol 1e-13 (run_nonlinear_matrixfree_split_mp.py:57-80,
    run_base.py:689-690); its additive constant never reaches f because
    h1 subtracts the mean (Toolbox.pyx:50-59).  The oracle solves exactly
    in Fourier space and returns the mean-free phi.
    """
    nx = b.shape[0]
    lam = poisson_eigenvalues(nx, hx, laplace)
    bh = np.fft.fft(b)
    ph = np.zeros_like(bh)
    ph[1:] = bh[1:] / lam[1:]
    return np.fft.ifft(ph).real


# --------------------------------------------------------------------------
# invariants            core/distribution_function.pyx:130-199,
#                       core/hamiltonian.pyx:67-100, core/potential.pyx:48-59
# --------------------------------------------------------------------------
def invariants(f, h0, h1, h2, phi, grid, charge):
    nx, nv = grid.nx, grid.nv
    fp = np.roll(f, -1, axis=0)
    mass = (f[:, :-1] + fp[:, :-1] + fp[:, 1:] + f[:, 1:]).sum() * 0.25 * grid.hx * grid.hv
    l2 = (grid.hx * grid.hv) * (f[:, :-1] ** 2).sum()
    ekin = (f * h0[None, :]).sum() * grid.hx * grid.hv
    epot = ((f * h1[:, None]).sum() + (f * h2[:, None]).sum()) * grid.hx * grid.hv
    energy = ekin + 0.5 * epot
    efield = 0.5 * charge * ((np.roll(phi, -1) - phi) ** 2).sum() / grid.hx
    return dict(mass=mass, l2=l2, energy=energy, ekin=ekin, epot=epot, efield=efield)


# --------------------------------------------------------------------------
# GMRES with PETSc KSPGMRES defaults (third-party, not vendored): restart 30,
# classical Gram-Schmidt without refinement, Givens rotations, zero initial
# guess, convergence when ||r_k|| <= max(rtol*||b||, atol), divergence when
# ||r_k|| >= dtol*||b||; called from SNESSolve_KSPONLY (ksp configured at
# run_nonlinear_matrixfree_split_mp.py:47-53, tolerances run_base.py:189-196).
# Textbook GMRES(m) of Saad & Schultz (1986) with those conventions.
# --------------------------------------------------------------------------
KSP_CONVERGED_RTOL = 2
KSP_CONVERGED_ATOL = 3
KSP_DIVERGED_ITS = -3
KSP_DIVERGED_DTOL = -4
KSP_DIVERGED_BREAKDOWN = -5
KSP_DIVERGED_NANORINF = -9


def gmres(apply_A, b, rtol=1e-5, atol=1e-50, max_it=10000, restart=30, dtol=1e5,
          haptol=1e-30, history=None):
    """Solve A x = b from x0 = 0.  Returns (x, its, reason)."""
    x = np.zeros_like(b)
    its = 0
    reason = 0
    rnorm0 = None
    ttol = None
    first = True
    def converged(res):
        if not np.isfinite(res):
            return KSP_DIVERGED_NANORINF
        if res <= ttol:
            return KSP_CONVERGED_ATOL if res < atol else KSP_CONVERGED_RTOL
        if res >= dtol * rnorm0:
            return KSP_DIVERGED_DTOL
        return 0

    while True:
        # initial residual of the cycle: b for the zero guess, b - A x after a restart
        r = b.copy() if first else b - apply_A(x)
        beta = math.sqrt(float(np.vdot(r, r).real))
        if history is not None:
            history.append(beta)
        if beta == 0.0:
            reason = KSP_CONVERGED_ATOL
            break
        if first:
            rnorm0 = beta
            ttol = max(rtol * rnorm0, atol)
        reason = converged(beta)
        if reason != 0:
            break
        first = False
        V = [r * (1.0 / beta)]
        H = np.zeros((restart + 1, restart))
        cs = np.zeros(restart)
        sn = np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        it = 0
        res = beta
        while reason == 0 and it < restart and its < max_it:
            w = apply_A(V[it])
            h = np.array([np.vdot(vj, w).real for vj in V])     # CGS: all dots first
            for jj in range(it + 1):
                w = w - h[jj] * V[jj]
            H[:it + 1, it] = h
            tt = math.sqrt(float(np.vdot(w, w).real))
            H[it + 1, it] = tt
            hapbnd = min(abs(tt / g[it]) if g[it] != 0 else 0.0, haptol)
            hapend = tt < hapbnd
            V.append(w * (1.0 / tt) if tt != 0.0 else w)
            for jj in range(it):        # previous rotations
                t = H[jj, it]
                H[jj, it] = cs[jj] * t + sn[jj] * H[jj + 1, it]
                H[jj + 1, it] = cs[jj] * H[jj + 1, it] - sn[jj] * t
            if not hapend:
                d = math.sqrt(H[it, it] ** 2 + H[it + 1, it] ** 2)
                cs[it] = H[it, it] / d
                sn[it] = H[it + 1, it] / d
                g[it + 1] = -(sn[it] * g[it])
                g[it] = cs[it] * g[it]
                H[it, it] = cs[it] * H[it, it] + sn[it] * H[it + 1, it]
                res = abs(g[it + 1])
            else:
                res = 0.0
            it += 1
            its += 1
            if history is not None:
                history.append(res)
            reason = converged(res)
            if hapend and reason == 0:
                reason = KSP_DIVERGED_BREAKDOWN
        # build the solution of this cycle: back substitution on the rotated H
        if it > 0:
            y = np.zeros(it)
            for k in range(it - 1, -1, -1):
                t = g[k]
                for jj in range(k + 1, it):
                    t -= H[k, jj] * y[jj]
                y[k] = t / H[k, k]
            upd = np.zeros_like(b)
            for jj in range(it):
                upd = upd + y[jj] * V[jj]
            x = x + upd
        if reason != 0:
            break
        if its >= max_it:
            reason = KSP_DIVERGED_ITS
            break
    return x, its, reason


# --------------------------------------------------------------------------
# the timestep driver       run/run_nonlinear_matrixfree_split_mp.py:111-174
#                           run/run_base.py:408-506, 623-763
# --------------------------------------------------------------------------
def maxwellian(v, vth=1.0, vd=0.0):
    return np.exp(-0.5 * ((v - vd) / vth) ** 2) / (math.sqrt(2. * np.pi) * vth)


def initial_distribution(grid, profile):
    """f_ij = profile(x_i, v_j) on rows 2..nv-3, 0 elsewhere (Toolbox.pyx:190-203)."""
    X, V = np.meshgrid(grid.x, grid.v, indexing='ij')
    f = np.asarray(profile(X, V), dtype=np.float64)
    f[:, :STENCIL] = 0.0
    f[:, grid.nv - STENCIL:] = 0.0
    return f


def normalise(f, grid):
    """nave = sum(f)*hv/nx; f *= 1/nave  (run_base.py:623-628)."""
    nave = f.T.reshape(-1).sum() * grid.hv / grid.nx
    return f * (1. / nave)


class OracleRun:
    """Picard/Newton-Krylov timestep loop of the reference run script.

    Mirrors run_nonlinear_matrixfree_split_mp.run() (:111-174): history,
    H update, pre-loop residual, Picard iterations of SNES(ksponly) +
    GMRES(P^-1 J) on P^-1 F, moments + Poisson after each solve, stop when
    the residual rises (reverting f), drops below snes_atol or after
    snes_max_iter iterations, abort when it exceeds 10.
    """

    def __init__(self, grid, f0, bracket='arakawaj4', laplace='cfd4', charge=-1.0,
                 mass=1.0, ksp_rtol=1e-10, ksp_atol=1e-50, ksp_max_it=1000,
                 snes_atol=1e-10, snes_max_iter=20, restart=30, precond_variant='lapack',
                 preconditioner='tensor_kinetic', dissipation=None, initial_guess=None, initial_iter=1,
                 external=None):
        self.grid = grid
        self.kind = bracket
        self.external = external
        self.itime = 0
        self.diss = dissipation
        if initial_guess not in (None, 'None', '') + GUESSES:
            raise ValueError("initial guess %r" % (initial_guess,))
        self.guess = initial_guess if initial_guess not in (None, 'None', '') else None
        self.niter = int(initial_iter)
        self.H = np.zeros((grid.nx, grid.nv))      # the solver's Have starts as a zero Vec
        self.laplace = laplace
        self.charge = charge
        self.ksp = dict(rtol=ksp_rtol, atol=ksp_atol, max_it=ksp_max_it, restart=restart)
        self.snes_atol = snes_atol
        self.snes_max_iter = snes_max_iter
        self.h0 = kinetic_hamiltonian(grid, mass)
        self.fc = normalise(np.array(f0, dtype=np.float64), grid)
        self.h2c = np.zeros(grid.nx)
        self.h2h = np.zeros(grid.nx)
        self.pc_ext = np.zeros(grid.nx)
        self.calculate_moments()
        self.calculate_external(0.)             # run_base.py:514-525
        # TensorProductPreconditioner.create (TensorProduct.pyx:66-86), built
        # after the initial moments as in run_base.py (its phi is pc_int)
        if preconditioner in ('tensor_kinetic', 'tensor_kinetic_scipy'):
            self.pc = Preconditioner(grid, precond_variant)
        elif preconditioner in ('tensor_potential', 'tensor_potential_scipy'):
            self.pc = PotentialPreconditioner(grid, self.phi, precond_variant)
        elif preconditioner in (None, 'None', 'none'):
            self.pc = IdentityPreconditioner()        # preconditioner_type None: KSP on J itself
        else:
            raise ValueError(preconditioner)
        self.log = []

    # --- moments / Poisson (run_base.py:640-704, 741-742)
    def calculate_moments(self, f=None):
        """n, u, e, a of f (run_base.py:640-651), then Poisson -> phi, h1."""
        f = self.fc if f is None else f
        self.n = density(f, self.grid)
        if self.diss is not None and self.diss.collisions:
            self.u = velocity_density(f, self.grid)
            self.e = energy_density(f, self.grid)
            self.a = collision_factor(self.n, self.u, self.e)
        b = poisson_rhs(self.n, self.charge, self.grid.nx)
        self.phi = poisson_solve(b, self.grid.hx, self.laplace)
        self.h1c = potential_to_hamiltonian(self.phi, self.grid.nx)

    def update_previous(self):
        self.H = h_average(self.h0, self.h1c, self.h1h, self.h2c, self.h2h)

    def calculate_external(self, t):
        """run_base.py:707-725: p_i = external(x_i, t) per point, the average
        removed (VecSum in index order, VecShift), then h2 =
        potential_to_hamiltonian(p) (copy_pext_to_h, :746-747).  No external
        potential: p and h2 stay zero."""
        if self.external is None:
            return
        p = np.array([float(self.external(x, t)) for x in self.grid.x])
        psum = 0.0
        for v in p:
            psum += v
        p = p + (-(psum / self.grid.nx))
        self.pc_ext = p
        self.h2c = potential_to_hamiltonian(p, self.grid.nx)

    def _mom(self):
        if not self.diss.collisions:
            return None
        return dict(np=self.n, up=self.u, ap=self.a, nh=self.nh, uh=self.uh, ah=self.ah)

    def _residual(self, f):
        if self.diss is None:
            return residual(self.kind, f, self.fh, self.H, self.grid)
        # function_solver also sets the average the Jacobian's double bracket uses
        self.fave = 0.5 * self.fh + 0.5 * f
        return residual_diss(self.kind, f, self.fh, self.H, self.grid, self.diss, self._mom())

    def function(self, f):
        return self.pc.apply(self._residual(f))

    def jacobian_apply(self, u):
        if self.diss is None:
            return self.pc.apply(jacobian(self.kind, u, self.H, self.grid))
        return self.pc.apply(jacobian_diss(self.kind, u, self.fave, self.H, self.grid, self.diss, self._mom()))

    def calculate_residual(self):
        """||F(fc)||_2 + ||A phi + (n - n_mean) q||_2  (run_base.py:631-637)."""
        fn = np.linalg.norm(self._residual(self.fc))
        pr = poisson_function_mult(self.phi, self.n, self.charge, self.grid.hx2_inv, self.laplace)
        return fn + np.linalg.norm(pr)

    def snes_ksponly(self):
        """SNESSolve_KSPONLY: F = P^-1 F(fc); KSPSolve; fc -= delta if the
        linear solve did not diverge."""
        b = self.function(self.fc)
        delta, its, reason = gmres(self.jacobian_apply, b, self.ksp['rtol'], self.ksp['atol'],
                                   self.ksp['max_it'], self.ksp['restart'])
        if reason > 0:
            self.fc = self.fc - delta
        return its, reason

    # --- explicit initial guess (run_base.py:786-927)
    def _guess_iteration(self):
        g, k = self.grid, self.kind
        h0 = np.broadcast_to(self.h0[None, :], (g.nx, g.nv))
        h12 = np.broadcast_to((self.h1h + self.h2h)[:, None], (g.nx, g.nv))
        if self.guess == 'symplectic2':
            self.fc = symplectic_kinetic(k, self.fc, h0, 0.5, g, self.niter)
            self.fc = symplectic_potential(k, self.fc, h12, 1.0, g, self.niter)
            self.fc = symplectic_kinetic(k, self.fc, h0, 0.5, g, self.niter)
        elif self.guess == 'symplectic4':
            c, d = symplectic4_coefficients()
            for q in range(3):
                self.fc = symplectic_kinetic(k, self.fc, h0, c[q], g, self.niter)
                self.fc = symplectic_potential(k, self.fc, h12, d[q], g, self.niter)
            self.fc = symplectic_kinetic(k, self.fc, h0, c[3], g, self.niter)
        else:
            # the solver's H = H0 + H1 + H2 of 2D Vecs (PETScArakawaRungeKutta.pyx:28-32)
            h = (h0 + self.h1h[:, None]) + self.h2h[:, None]
            rk = rk4_16 if self.guess == 'rk4_16' else rk4_18   # 'rk4' -> rk18 (run_base.py:825)
            self.fc = rk(k, self.fc, h, g, self.niter)
        self.calculate_moments()

    def initial_guess(self):
        """run_base.py:786-805: back up f, take niter explicit iterations
        (moments after each), keep them only if the residual went down.  As in
        the reference, a reverted guess leaves the guessed moments in place."""
        fl = self.fc.copy()
        prev = self.calculate_residual()
        for _ in range(self.niter):
            self._guess_iteration()
        ig = self.calculate_residual()
        if ig > prev:
            self.fc = fl
        return prev, ig

    def step(self):
        # make_history (run_base.py:749-763) + update_history (PETScVlasovSolver.pyx:135-139)
        self.fh = self.fc.copy()
        self.h1h = self.h1c.copy()
        self.h2h = self.h2c.copy()
        if self.diss is not None and self.diss.collisions:
            self.nh, self.uh, self.ah = self.n.copy(), self.u.copy(), self.a.copy()
        # calculate_external(current_time) (run_nonlinear_matrixfree_split_mp.py:113, 124-125)
        self.itime += 1
        self.calculate_external(self.grid.ht * self.itime)
        # update_history -> preconditioner.update_matrices (PETScVlasovSolver.pyx:135-139)
        if isinstance(self.pc, PotentialPreconditioner):
            self.pc.update(self.phi)
        guess = self.initial_guess() if self.guess else None
        self.update_previous()
        pred = self.calculate_residual()
        entry = dict(initial=pred, picard=[])
        if guess is not None:
            entry['guess'] = guess
        i = 0
        while True:
            i += 1
            fl = self.fc.copy()
            its, reason = self.snes_ksponly()
            self.calculate_moments()
            self.update_previous()
            prev, pred = pred, self.calculate_residual()
            entry['picard'].append(dict(gmres_its=its, reason=reason, residual=pred))
            if (pred > prev and i > 1) or pred < self.snes_atol or i >= self.snes_max_iter:
                if pred > prev:
                    self.fc = fl
                    self.calculate_moments()
                break
        self.log.append(entry)
        if pred > 10.:
            raise RuntimeError("ERROR: Residual of nonlinear solver too large.")
        return entry
