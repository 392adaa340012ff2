"""Grid, configuration and initial data for the implicit Vlasov-Poisson step.

Mirrors the parts of the reference the hot path depends on:
``vlasov/core/Grid.pyx:14-51`` (grid constants),
``vlasov/core/config.py:50-167`` with ``runspec.cfg:1-134`` (the option
subset that selects the solver), ``vlasov/run/run_base.py:118-137, 408-506,
623-628`` (grid set-up, initial data, normalisation) and
``vlasov/solvers/preconditioner/TensorProduct.pyx:109-116`` (kinetic
eigenvalues).
"""

import importlib
import math
import os

import numpy as np

STENCIL = 2   # run_base.py:48-49


class Grid:
    """Grid constants (Grid.pyx:14-51).

    hx = L/nx (periodic x, x_i = i hx), hv = (vmax - vmin)/(nv - 1)
    (v_j = vmin + j hv), as run_base.py:130-137 and the DMDA uniform
    coordinates of run_base.py:252-254 define them.
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
        self.t = np.linspace(0, self.ht * self.nt, self.nt + 1)
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

    def xLength(self):
        return self.L

    def __repr__(self):
        return "Grid(nx=%d, nv=%d, L=%g, v=[%g,%g], ht=%g)" % (self.nx, self.nv, self.L, self.vmin, self.vmax, self.ht)


def kinetic_eigenvalues(n):
    """Eigenvalues of the periodic x-difference used by the kinetic
    preconditioner, evaluated exactly as TensorProduct.compute_eigenvalues
    (TensorProduct.pyx:109-116) does, so the device receives the same bits."""
    eig = np.empty(n, dtype=np.complex128)
    for i in range(n):
        eig[i] = np.exp(2.j * np.pi * float(i) / n * (n - 1)) - np.exp(2.j * np.pi * float(i) / n)
    return np.fft.ifftshift(eig)


def maxwellian(v, vth=1.0, vd=0.0):
    return np.exp(-0.5 * ((v - vd) / vth) ** 2) / (math.sqrt(2. * np.pi) * vth)


# ---------------------------------------------------------------------------
# configuration: the hot-path subset of runspec.cfg with its defaults
# ---------------------------------------------------------------------------
RUNSPEC_DEFAULTS = {
    'grid': {'nt': 100, 'nx': 128, 'nv': 256, 'ht': 0.1, 'L': 0.0, 'vmin': None, 'vmax': 5.0},
    'solver': {'method': 'iterative', 'type': 'nonlinear', 'mode': 'split', 'timestepping': 'mp',
               'poisson_bracket': 'ArakawaJ4', 'laplace_operator': 'CFD4', 'averaging_operator': None,
               'dissipation_type': None, 'collision_operator': None, 'bracket_operator': None,
               'preconditioner_type': None, 'preconditioner_scheme': None, 'initial_guess': None,
               'lu_package': 'mumps', 'initial_iter': 1,
               'petsc_ksp_rtol': 1E-4, 'petsc_ksp_atol': 1E-50, 'petsc_ksp_max_iter': 1000,
               'petsc_snes_atol': 1E-10, 'petsc_snes_rtol': 1E-14, 'petsc_snes_stol': 1E-16,
               'petsc_snes_max_iter': 20, 'coll_freq': 0.0, 'coll_diff': 1.0, 'coll_drag': 1.0,
               'regularisation': 0.0},
    'initial_data': {'mass': 1.0, 'charge': 1.0, 'density': 1.0, 'temperature': 1.0,
                     'distribution_python': None, 'density_python': None, 'temperature_python': None,
                     'potential_python': None, 'external_python': None, 'nfourier': 0},
    'io': {'output_dir': '.', 'nsave': 1, 'hdf5_input': None, 'hdf5_output': None},
}

_INT_KEYS = {'nt', 'nx', 'nv', 'initial_iter', 'petsc_ksp_max_iter', 'petsc_snes_max_iter', 'nfourier', 'nsave'}


def _convert(key, value):
    if isinstance(value, str):
        s = value.strip().strip("'\"")
        if s.lower() == 'none' or s == '':
            return None
        if key in _INT_KEYS:
            return int(float(s))
        try:
            return float(s)
        except ValueError:
            return s
    return value


def parse_cfg(text):
    """Parse a viVlasov1D configuration file (configobj INI subset)."""
    out = {}
    section = None
    for raw in text.splitlines():
        line = raw.split('#', 1)[0].strip()
        if not line:
            continue
        if line.startswith('[') and line.endswith(']'):
            section = line[1:-1].strip()
            out.setdefault(section, {})
            continue
        if '=' in line and section is not None:
            k, v = line.split('=', 1)
            out[section][k.strip()] = v.strip()
    return out


class Config(dict):
    """Run configuration with the getters of vlasov/core/config.py:50-167."""

    def __init__(self, infile=None, **sections):
        super().__init__()
        for sec, vals in RUNSPEC_DEFAULTS.items():
            self[sec] = dict(vals)
        data = {}
        if isinstance(infile, dict):
            data = infile
        elif isinstance(infile, str):
            with open(infile) as fh:
                data = parse_cfg(fh.read())
        for src in (data, sections):
            for sec, vals in src.items():
                self.setdefault(sec, {})
                for k, v in vals.items():
                    self[sec][k] = _convert(k, v)

    @staticmethod
    def _lower(v):
        return v.lower() if isinstance(v, str) else v

    def get_timestepping(self):
        return self['solver']['timestepping'].lower()

    def get_poisson_bracket(self):
        return self['solver']['poisson_bracket'].lower()

    def get_laplace_operator(self):
        return self['solver']['laplace_operator'].lower()

    def is_averaging_operator_none(self):
        v = self['solver']['averaging_operator']
        return v is None or v.lower() == 'none'

    def get_averaging_operator(self):
        return "point" if self.is_averaging_operator_none() else self['solver']['averaging_operator'].lower()

    def is_preconditioner_none(self):
        t, s = self['solver']['preconditioner_type'], self['solver']['preconditioner_scheme']
        return t is None or t.lower() == 'none' or s is None or s.lower() == 'none'

    def get_preconditioner(self):
        return '' if self.is_preconditioner_none() else self['solver']['preconditioner_scheme'].lower()

    def is_dissipation_none(self):
        d = self['solver']['dissipation_type']
        if d is None or self['solver']['coll_freq'] == 0.:
            return True
        return d.lower() == 'none'

    def get_collision_operator(self):
        d = self['solver']['dissipation_type']
        if d is None or self['solver']['coll_freq'] == 0. or d.lower() != 'collisions':
            return ''
        return self['solver']['collision_operator'].lower()

    def get_double_bracket(self):
        d = self['solver']['dissipation_type']
        if d is None or self['solver']['coll_freq'] == 0. or d.lower() != 'double_bracket':
            return ''
        return self['solver']['bracket_operator'].lower()

    def is_regularisation_none(self):
        return self['solver']['regularisation'] == 0.

    def make_grid(self):
        g = self['grid']
        return Grid(g['nx'], g['nv'], g['L'], g['vmax'], g['vmin'], g['ht'], g['nt'])


# ---------------------------------------------------------------------------
# initial data                         Toolbox.pyx:175-224; run_base.py:623-628
# ---------------------------------------------------------------------------
def initial_distribution(grid, profile):
    """f_ij = profile(x_i, v_j) on rows 2..nv-3, zero on the four v-boundary
    rows (Toolbox.pyx:190-203).  Returned in natural order, shape (nv, nx)."""
    X, V = np.meshgrid(grid.x, grid.v, indexing='xy')      # (nv, nx)
    f = np.asarray(profile(X, V), dtype=np.float64)
    f = np.broadcast_to(f, (grid.nv, grid.nx)).copy()
    f[:STENCIL, :] = 0.0
    f[grid.nv - STENCIL:, :] = 0.0
    return f


def initial_distribution_nT(grid, n, T):
    """Maxwellian with density n(x), temperature T(x) (Toolbox.pyx:206-224)."""
    fac = math.sqrt(0.5 / np.pi)
    V = grid.v[:, None]
    f = n[None, :] * fac * np.exp(-0.5 * V ** 2 / T[None, :]) / np.sqrt(T[None, :])
    f[:STENCIL, :] = 0.0
    f[grid.nv - STENCIL:, :] = 0.0
    return f


def fourier_filter(f, nfourier):
    """The initial-data Fourier filter along x (run_base.py:467-496):
    scipy.fftpack.rfft over x, entries nfourier and up set to zero, irfft.
    FFTPACK stores the coefficients as y0, Re y1, Im y1, Re y2, Im y2, ...
    (Re y_{nx/2} last for even nx), so Re y_k survives when 2k - 1 < nfourier
    and Im y_k when 2k < nfourier.  f in natural order, shape (nv, nx)."""
    if nfourier <= 0:
        return f
    nx = f.shape[1]
    Y = np.fft.rfft(f, axis=1)
    k = np.arange(Y.shape[1])
    keep_re = np.where(k == 0, True, 2 * k - 1 < nfourier)
    keep_im = np.where(k == 0, True, 2 * k < nfourier)
    Y = np.where(keep_re, Y.real, 0.0) + 1j * np.where(keep_im, Y.imag, 0.0)
    return np.fft.irfft(Y, n=nx, axis=1)


def initial_function(name, attr):
    """The function `attr` of the initial-data module `name`: the reference
    imports runs.<name> (run_base.py:418, 433, 445); a module of that name on
    sys.path is taken as well."""
    for modname in ("runs." + name, name):
        try:
            mod = importlib.import_module(modname)
        except ImportError:
            continue
        if hasattr(mod, attr):
            return getattr(mod, attr)
    raise ImportError("initial-data module %r with a function %r not found (runs.%s or %s)"
                      % (name, attr, name, name))


def initial_data(cfg, grid, profile=None):
    """f0 as run_base.py:408-496 builds it, before normalisation: a profile
    (here a callable, or the config's distribution_python module's
    distribution(x, v), evaluated per point as Toolbox.initialise_distribution_function
    does), else a Maxwellian in n(x), T(x) from density_python /
    temperature_python (density(x_i, L), temperature(x_i)) or the constant
    density / temperature; then the nfourier filter.  Natural order (nv, nx)."""
    idata = cfg['initial_data']
    if profile is None and idata.get('distribution_python'):
        dist = initial_function(idata['distribution_python'], 'distribution')
        profile = np.vectorize(lambda x, v: float(dist(x, v)), otypes=[np.float64])
    if profile is not None:
        f = initial_distribution(grid, profile)
    else:
        if idata.get('density_python'):
            dens = initial_function(idata['density_python'], 'density')
            n = np.array([dens(x, grid.L) for x in grid.x], dtype=np.float64)
        else:
            n = np.full(grid.nx, float(idata['density']))
        if idata.get('temperature_python'):
            temp = initial_function(idata['temperature_python'], 'temperature')
            T = np.array([temp(x) for x in grid.x], dtype=np.float64)
        else:
            T = np.full(grid.nx, float(idata['temperature']))
        f = initial_distribution_nT(grid, n, T)
    return fourier_filter(f, int(idata.get('nfourier') or 0))


def normalise(f, grid):
    """nave = sum(f) hv / nx; f *= 1/nave (run_base.py:623-628)."""
    nave = f.reshape(-1).sum() * grid.hv / grid.nx
    return f * (1. / nave)


def kinetic_hamiltonian(grid, mass=1.0):
    """h0_j = 0.5 v_j^2 mass (Toolbox.pyx:175-187), the v-only factor."""
    return 0.5 * grid.v ** 2 * mass


# profiles of the benchmark workloads ---------------------------------------
def landau(k, amp):
    return lambda x, v: (1. + amp * np.cos(k * x)) * maxwellian(v)


def twostream(k, amp, vd=2.4):
    return lambda x, v: (1. + amp * np.cos(k * x)) * 0.5 * (maxwellian(v, 1.0, vd) + maxwellian(v, 1.0, -vd))


def bump_on_tail(k, amp):
    def prof(x, v):
        base = 0.9 * maxwellian(v) + 0.1 * maxwellian((v - 4.5) / 0.5) / 0.5
        return (1. + amp * np.cos(k * x)) * base
    return prof
