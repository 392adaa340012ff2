"""Time loop of the split-mode, matrix-free, nonlinear implicit-midpoint run.

``Runscript`` mirrors ``viVlasov1Drunscript`` of
vlasov/run/run_nonlinear_matrixfree_split_mp.py:15-174 together with the
parts of ``viVlasov1Dbase`` (vlasov/run/run_base.py:35-807) the timestep
uses: initial data and normalisation, moments + Poisson, residual, history,
and the Picard loop with its stop / revert / abort rules, the collision /
double-bracket dissipation terms and the explicit initial guess the
configuration selects, and (optional) HDF5 output in the reference's schema.  The other run modes are not part of
this library.
"""

import json
import sys
import time

import numpy as np
import torch

from . import _lib
from .core import Config, Grid, initial_data, initial_function, kinetic_hamiltonian, normalise
from .solvers import (Context, PETScArakawaRungeKutta, PETScArakawaSymplectic, PETScPoissonSolver,
                      PETScVlasovSolver, SNES, Toolbox, effective_preconditioner)


class ResidualTooLarge(RuntimeError):
    """Raised where the reference exits: residual > 10 after a step (:165-168)."""


class Runscript:
    def __init__(self, cfg, profile=None, f0=None, log=None, restart=30, verbose=False, precision='double',
                 inner_rtol=1e-5, normalise_f0=True, refine='newton', hdf5_out=None, external=None,
                 fp64_above='auto'):
        if not isinstance(cfg, Config):
            cfg = Config(cfg)
        self.cfg = cfg
        self.grid = grid = cfg.make_grid()
        self.verbose = verbose
        self.charge = cfg['initial_data']['charge']
        self.mass = cfg['initial_data']['mass']
        sol = cfg['solver']
        self.ctx = Context(grid, bracket=cfg.get_poisson_bracket(),
                           preconditioner=effective_preconditioner(cfg.get_preconditioner()),
                           laplace=cfg.get_laplace_operator(), charge=self.charge, restart=restart)
        ctx = self.ctx
        # vectors (run_base.py:192-233); x-vectors are 1D, the Hamiltonians are their factors
        self.h0 = torch.from_numpy(kinetic_hamiltonian(grid, self.mass)).cuda()
        self.h1c, self.h1h = ctx.new_x(), ctx.new_x()
        self.h2c, self.h2h = ctx.new_x(), ctx.new_x()
        self.fc, self.fh, self.fl, self.fb = ctx.new_field(), ctx.new_field(), ctx.new_field(), ctx.new_field()
        self.N = ctx.new_x()
        # velocity / energy density and collision factor, current and history
        # (run_base.py:313-320); refreshed each Picard iteration only when a
        # collision operator reads them
        self.U, self.E, self.A = ctx.new_x(), ctx.new_x(), ctx.new_x()
        self.Nh, self.Uh, self.Eh, self.Ah = ctx.new_x(), ctx.new_x(), ctx.new_x(), ctx.new_x()
        self.collisions = bool(cfg.get_collision_operator())
        self.pc_int, self.ph_int = ctx.new_x(), ctx.new_x()
        self.pc_ext = ctx.new_x()
        self.toolbox = Toolbox(None, None, grid, ctx)
        # initial data (run_base.py:408-506): f0 as given, else the profile
        # callable or the configuration's initial-data settings
        if f0 is None:
            f0 = initial_data(cfg, grid, profile)
        f0 = np.asarray(f0, dtype=np.float64).reshape(grid.nv, grid.nx)
        if normalise_f0:
            f0 = normalise(f0, grid)
        self.fc.copy_(torch.from_numpy(f0))
        self.calculate_moments()
        # external potential external(x, t) (run_base.py:514-525): the callable
        # given, else the configuration's external_python module
        ext = cfg['initial_data'].get('external_python')
        self.external = external if external is not None else (initial_function(ext, 'external') if ext else None)
        self.calculate_external(0.)
        # solver objects (run_nonlinear_matrixfree_split_mp.py:26-80)
        self.vlasov_solver = PETScVlasovSolver(cfg, None, None, grid, self.h0, self.h1c, self.h1h, self.h2c,
                                               self.h2h, self.pc_int, self.charge, coll_freq=sol['coll_freq'],
                                               coll_diff=sol['coll_diff'], coll_drag=sol['coll_drag'],
                                               regularisation=sol.get('regularisation', 0.0), context=ctx)
        self.vlasov_solver.set_moments(self.N, self.U, self.E, self.A, self.Nh, self.Uh, self.Eh, self.Ah)
        self.poisson_solver = PETScPoissonSolver(None, grid.nx, grid.hx, self.charge,
                                                 laplace=cfg.get_laplace_operator(), context=ctx)
        self.snes = SNES()
        self.snes.setType('ksponly')
        self.snes.setFunction(self.vlasov_solver.function_snes_mult, self.fb)
        self.snes.getKSP().setTolerances(rtol=sol['petsc_ksp_rtol'], atol=sol['petsc_ksp_atol'],
                                         max_it=sol['petsc_ksp_max_iter'])
        # 'mixed': FP32 GMRES cycles; the residual, update, moments and
        # Poisson solve stay FP64.  refine='newton': the Picard loop below is
        # the refinement (one FP32 cycle per solve, to inner_rtol, until the
        # FP64 residual meets snes_atol); 'ksp': each solve refined in FP64 to
        # petsc_ksp_rtol.  fp64_above (with 'newton'): a Picard solve that
        # starts from a residual above it runs in FP64 at the same relative
        # tolerance; 'auto' = snes_atol * 2^20: the FP32 rounding of a
        # correction (2^-24 of it, the correction being about the size of the
        # residual it removes) then stays 16 times below the nonlinear
        # tolerance; None: every solve FP32
        self.snes_atol = sol['petsc_snes_atol']
        if fp64_above == 'auto':
            fp64_above = self.snes_atol * 2.0 ** 20
        self.snes.getKSP().setPrecision(precision, inner_rtol=inner_rtol, refine=refine, fp64_above=fp64_above)
        self.precision = precision
        self.snes_max_iter = sol['petsc_snes_max_iter']
        # initial guess (run_base.py:148, 171-183, 368-372): the steppers see the
        # history potentials h1h, h2h; 'rk4' is the 3/8 rule the reference's
        # initial_guess_rk4 names (rk18, :825)
        self.nInitial = int(sol.get('initial_iter', 1) or 1)
        self.arakawa_rk4 = PETScArakawaRungeKutta(cfg, None, grid, self.h0, self.h1h, self.h2h, self.nInitial,
                                                  context=ctx)
        self.arakawa_symplectic = PETScArakawaSymplectic(cfg, None, grid, self.h0, self.h1h, self.h2h,
                                                         self.nInitial, context=ctx)
        guesses = {None: None, 'None': None, '': None, 'rk4': self.initial_guess_rk4,
                   'rk4_18': self.initial_guess_rk4, 'rk4_16': self.initial_guess_rk4_16,
                   'symplectic2': self.initial_guess_symplectic2, 'symplectic4': self.initial_guess_symplectic4}
        ig = sol.get('initial_guess')
        if ig == 'gear':
            # run_base.py:510-511 calls PETScArakawaGear.initialise_history,
            # which is commented out (PETScArakawaGear.pyx:70-103): the
            # reference stops at set-up with this option
            raise _lib.VVError("initial guess 'gear' is not provided: the reference's Gear history set-up "
                               "(PETScArakawaGear.initialise_history) does not exist")
        if ig not in guesses:
            raise _lib.VVError("initial guess %r is not provided (rk4, rk4_16, rk4_18, symplectic2, symplectic4)"
                               % (ig,))
        self.initial_guess_method = guesses[ig]
        self.itime = 0
        self.history = []
        self.log = log
        # HDF5 output in the reference's schema (run_base.py:528-562, 766-783):
        # the initial state as timestep 0, then every nsave'th step
        self.nsave = max(1, int(cfg['io'].get('nsave', 1) or 1))
        path = hdf5_out if hdf5_out is not None else cfg['io'].get('hdf5_output')
        self.output = None
        if path:
            from .hdf5 import HDF5Output
            self.output = HDF5Output(path, grid.x, grid.v, self.charge)
            self.save_hdf5_vectors()

    # run_base.py:631-637 --------------------------------------------------
    def calculate_residual(self):
        """||F(fc)||_2 + ||A phi + (n - mean n) q||_2.  The Poisson part is the
        residual of the last moments/Poisson solve, which the fused moments
        kernel leaves on the device (the reference recomputes it from the same
        phi, n); both norms come back in one read."""
        fnorm, pnorm = self.ctx.residual_norms(self.fc, self.vlasov_solver.Fh, self.fb)
        return fnorm + pnorm

    # run_base.py:640-704, 741-742 -------------------------------------------
    def calculate_moments(self):
        self.ctx.moments(self.fc, self.N, self.pc_int, self.h1c, want_norm=False)
        if self.collisions:
            self.ctx.velocity_moments(self.fc, self.U, self.E)
            self.ctx.collision_factor(self.N, self.U, self.E, self.A)

    # run_base.py:707-725, 745-747 -------------------------------------------
    def calculate_external(self, t):
        """p_i = external(x_i, t), evaluated per point on the host as the
        reference does, its average removed (index-order sum), then
        h2 = potential_to_hamiltonian(p) on the device.  Without an external
        potential p and h2 stay zero."""
        if self.external is None:
            return
        g = self.grid
        p = np.array([float(self.external(x, t)) for x in g.x])
        psum = 0.0
        for v in p:
            psum += v
        p = p + (-(psum / g.nx))
        self.pc_ext.copy_(torch.from_numpy(p))
        self.ctx.potential_to_hamiltonian(self.pc_ext, self.h2c)

    # run_base.py:749-763 ----------------------------------------------------
    def make_history(self):
        self.fh.copy_(self.fc)
        self.h1h.copy_(self.h1c)
        self.h2h.copy_(self.h2c)
        self.ph_int.copy_(self.pc_int)
        if self.collisions:
            for cur, his in ((self.N, self.Nh), (self.U, self.Uh), (self.E, self.Eh), (self.A, self.Ah)):
                his.copy_(cur)
        self.vlasov_solver.update_history(self.fc)

    # run_base.py:786-927 --------------------------------------------------
    def initial_guess(self):
        """Back up f, take nInitial explicit iterations (moments after each)
        and keep them only if the residual went down.  As in the reference, an
        undone guess leaves its moments in place until the Picard loop's."""
        if self.initial_guess_method is None:
            return None
        self.fl.copy_(self.fc)
        prev = self.calculate_residual()
        self.initial_guess_method()
        ig = self.calculate_residual()
        if ig > prev:
            self.fc.copy_(self.fl)
        return prev, ig

    def initial_guess_rk4(self):
        for _ in range(self.nInitial):
            self.arakawa_rk4.rk4_18(self.fc)
            self.calculate_moments()

    def initial_guess_rk4_16(self):
        for _ in range(self.nInitial):
            self.arakawa_rk4.rk4_16(self.fc)
            self.calculate_moments()

    def initial_guess_symplectic2(self):
        for _ in range(self.nInitial):
            self.arakawa_symplectic.kinetic(self.fc, 0.5)
            self.arakawa_symplectic.potential(self.fc, 1.0)
            self.arakawa_symplectic.kinetic(self.fc, 0.5)
            self.calculate_moments()

    def initial_guess_symplectic4(self):
        fac2 = 2. ** (1. / 3.)
        c1 = 0.5 / (2. - fac2)
        c2 = c1 * (1. - fac2)
        d1 = 1. / (2. - fac2)
        d2 = - d1 * fac2
        for _ in range(self.nInitial):
            for c, d in ((c1, d1), (c2, d2), (c2, d1)):
                self.arakawa_symplectic.kinetic(self.fc, c)
                self.arakawa_symplectic.potential(self.fc, d)
            self.arakawa_symplectic.kinetic(self.fc, c1)
            self.calculate_moments()

    def step(self):
        """One timestep (run_nonlinear_matrixfree_split_mp.py:112-171)."""
        self.itime += 1
        self.make_history()
        self.calculate_external(self.itime * self.grid.ht)      # _mp.py:113, 124-125
        guess = self.initial_guess()
        self.vlasov_solver.update_previous(self.fc)
        pred = self.calculate_residual()
        rec = dict(step=self.itime, initial=pred, picard=[])
        if guess is not None:
            rec['guess'] = guess
        i = 0
        while True:
            i += 1
            self.fl.copy_(self.fc)
            self.snes.picard_residual = pred
            self.snes.solve(None, self.fc)
            self.snes.picard_residual = None
            self.calculate_moments()
            self.vlasov_solver.update_previous(self.fc)
            prev, pred = pred, self.calculate_residual()
            ksp = self.snes.getKSP()
            it_rec = dict(gmres=ksp.getIterationNumber(), reason=ksp.getConvergedReason(), residual=pred)
            if ksp.precision == 'mixed':
                it_rec['cycles'] = ksp.cycles
                it_rec['fp64'] = ksp.last_precision == 'double'
            rec['picard'].append(it_rec)
            if self.verbose:
                print("  Nonlinear Solver: %5i GMRES  iterations, residual = %24.16E"
                      % (ksp.getIterationNumber(), pred), flush=True)
            if (pred > prev and i > 1) or pred < self.snes_atol or i >= self.snes_max_iter:
                if pred > prev:
                    self.fc.copy_(self.fl)
                    self.calculate_moments()
                    rec['reverted'] = True
                break
        rec['final'] = pred
        self.history.append(rec)
        if self.log is not None:
            self.log.write(json.dumps(rec) + "\n")
        if pred > 10.:
            raise ResidualTooLarge("ERROR: Residual of nonlinear solver too large.")
        self.save_to_hdf5(self.itime)
        return rec

    # run_base.py:766-783 ----------------------------------------------------
    def save_to_hdf5(self, itime):
        if self.output is not None and itime % self.nsave == 0:
            self.save_hdf5_vectors()

    def save_hdf5_vectors(self):
        """t, h0, h1, h2, f, phi_int, phi_ext, n, u, e of the current state; the
        Hamiltonian parts are written as the reference's 2D fields (h0 v-only,
        h1 and h2 x-only, broadcast)."""
        g = self.grid
        self.ctx.velocity_moments(self.fc, self.U, self.E)
        h0 = np.broadcast_to(self.h0.cpu().numpy()[:, None], (g.nv, g.nx))
        h1 = np.broadcast_to(self.h1c.cpu().numpy()[None, :], (g.nv, g.nx))
        h2 = np.broadcast_to(self.h2c.cpu().numpy()[None, :], (g.nv, g.nx))
        self.output.save(self.itime * g.ht, h0, h1, h2, self.fc, self.pc_int, self.pc_ext, self.N, self.U, self.E)

    def close(self):
        if self.output is not None:
            self.output.close()
            self.output = None

    def set_state(self, f):
        """Replace the distribution function (a (nv, nx) float64 CUDA tensor
        or array) and recompute its moments, as after initialisation."""
        if isinstance(f, torch.Tensor):
            self.fc.copy_(f.reshape(self.grid.nv, self.grid.nx))
        else:
            self.fc.copy_(torch.from_numpy(np.asarray(f, dtype=np.float64).reshape(self.grid.nv, self.grid.nx)))
        self.calculate_moments()

    def run(self, nt=None):
        nt = self.grid.nt if nt is None else nt
        for _ in range(nt):
            self.step()

    def invariants(self):
        return self.ctx.invariants(self.fc, self.h0, self.h1c, self.h2c, self.pc_int)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.ctx.synchronize()
        self.close()
